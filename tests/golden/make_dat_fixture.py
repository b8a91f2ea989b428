"""Writes tests/golden/f1_text_mesh.npz: the inputs of the reference's
data/F1_text mesh (6,675 cells) as read by hydro_amd.dat (topology, node
coordinates, bed, initial state, Manning n, KLAS, boundary series for two
rows), so the GPU tests can run a real reference mesh where the reference
checkout is absent.  Run from the repository root with the checkout at
/root/reference."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "hydro-bench_amd"))


def main(src="/root/reference/data/F1_text", out=os.path.join(ROOT, "tests", "golden", "f1_text_mesh.npz")):
    import hydro_amd.dat as dat
    import hydro_amd.mesh as mesh

    captured = {}
    orig = mesh._finish

    def capture(name, raw, dtype, lib=None):
        captured.update(raw)
        return orig(name, raw, dtype, lib)

    dat._finish = capture
    try:
        dat.load_dat(src, ndays=2)
    finally:
        dat._finish = orig
    keep = {k: np.asarray(v) for k, v in captured.items()}
    np.savez_compressed(out, **keep)
    print(out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
