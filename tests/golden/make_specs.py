"""Write the mesh specs that the golden-vector script reads (run from the repo root)."""
import importlib.util
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "packaging-tm-modeling-tool_amd")


def _pkg():
    spec = importlib.util.spec_from_file_location("fcbga_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["fcbga_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


if __name__ == "__main__":
    _pkg()
    from fcbga_amd import cases
    for name, sp in [("small_box", cases.small_box_spec()), ("plate_16", cases.plate_spec(16)),
                     ("cfg1", cases.cfg1_spec()), ("cfg2", cases.cfg2_spec())]:
        with open(os.path.join(HERE, name + "_spec.json"), "w") as f:
            json.dump(sp, f, indent=1)
        print("wrote", name)
