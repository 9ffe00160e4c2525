"""Writes the call-program fixtures program_<scene>.json (run in the build
container only; the reference never travels).

For each reference scene file, the scene is built against the reference's
own scenelib (scenes/scenelib.py, imported from /root/reference/scenes), with
every public builder call on the Scene recorded as data (method name,
positional and keyword arguments; transform steps are the plain dicts the
step helpers return). The fixture holds that call log (the input vector)
and the backend call program the reference's Scene._program("scenes")
produces from it (scenes/scenelib.py:528-679, the output vector). Only data
is written: no text of the scene files or of scenelib.

    python tests/golden/make_program_fixtures.py
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SCENES = "/root/reference/scenes"
SCENES = ["smoke", "01-marble-run", "02-cornell-lume", "features", "06-spot-cascade", "09-ember-shore", "10-suncatcher", "14-toy-factory"]


def main():
    sys.path.insert(0, REF_SCENES)
    import scenelib as ref  # the reference builder (scenes/scenelib.py)

    log = []
    depth = [0]

    class Recording(ref.Scene):
        # records the scene file's own builder calls, not the calls the
        # builder makes on itself
        def __getattribute__(self, name):
            attr = object.__getattribute__(self, name)
            if name.startswith("_") or not callable(attr) or name in ("run",):
                return attr

            def call(*args, **kwargs):
                if depth[0] == 0:
                    log.append({"method": name, "args": list(args), "kwargs": kwargs})
                depth[0] += 1
                try:
                    return attr(*args, **kwargs)
                finally:
                    depth[0] -= 1
            return call

    ref.Scene = Recording
    for name in SCENES:
        del log[:]
        path = os.path.join(REF_SCENES, name + ".py")
        with open(path) as f:
            code = compile(f.read(), path, "exec")
        g = {"__name__": "sundog_scene_fixture", "__file__": path}
        exec(code, g)
        scene = g["s"]
        program = [list(c) for c in scene._program("scenes")]
        out = {"scene": "scenes/%s.py" % name, "builder_calls": log[:], "program": program}
        dst = os.path.join(HERE, "program_%s.json" % name)
        with open(dst, "w") as f:
            json.dump(out, f, separators=(",", ":"), default=lambda o: list(o) if isinstance(o, tuple) else repr(o))
        print("%s: %d builder calls, %d program calls -> %s" % (name, len(log), len(program), dst))


if __name__ == "__main__":
    main()
