"""Writes an input file of the C drop-in caller (tests/dropin_caller.c) with
the oracle's model setup, so that the caller can run outside pytest, for
instance under a profiler:

    python tests/dropin_caller_input.py OUT NT MODE [NX NY FDORDER FW FREE_SURF L]

(defaults 1200 300 8 20 1 2: the viscoelastic FDORDER 8 grid of the speed
test in test_gpu_dropin.py; MODE 0 copies, 1 resident, 2 resident with
deferred seismograms), then  tests/dropin_caller OUT result.bin."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))

from test_gpu_dropin import _caller_input  # noqa: E402


def main(argv):
    if len(argv) not in (4, 10):
        print(__doc__, file=sys.stderr)
        return 2
    out, nt, mode = argv[1], int(argv[2]), int(argv[3])
    nx, ny, fdorder, fw, free_surf, L = (1200, 300, 8, 20, 1, 2) if len(argv) == 4 else map(int, argv[4:10])
    _caller_input(out, nx, ny, fdorder, fw, free_surf, L, nt, mode)
    print("wrote %s: %dx%d, FDORDER %d, FW %d, FREE_SURF %d, L %d, %d steps, mode %d" % (
        out, nx, ny, fdorder, fw, free_surf, L, nt, mode))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
