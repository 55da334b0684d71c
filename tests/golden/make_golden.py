"""Writes tests/golden/reference_pulse_ik.npz from oracle/_ref/libref.so, the reference's own
tools/solver_common.c compiled from its sources (oracle/Makefile).  Fixtures: ik(i), imu(i) for
i < 2*65*64 (every angular index of up to 64 kappa shells) and beoyndDipolePulse_axialPart f_alpha(t)
for 200 seeded (E0, omega, cycles, t), stored as their exact bit patterns.

Run from the repository root after `make -C oracle`:  python3 tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libref.so is not built (needs the reference sources)")
    idx = np.arange(2 * 65 * 64, dtype=np.int32)
    ik = np.array([R.ik(int(i)) for i in idx], np.int32)
    imu = np.array([R.imu(int(i)) for i in idx], np.float64)

    class Pulse(C.Structure):
        _fields_ = [("E0", C.c_double), ("omega", C.c_double), ("T", C.c_double)]
    R.beyondDipolePulse_init.argtypes = [C.POINTER(Pulse), C.c_double, C.c_double, C.c_double]
    R.beoyndDipolePulse_axialPart.argtypes = [C.POINTER(Pulse), C.c_double, O.cplx_p]
    rng = np.random.default_rng(0x6ADE)
    args = np.stack([rng.uniform(1, 100, 200), rng.uniform(25, 75, 200), rng.integers(1, 16, 200).astype(float),
                     rng.uniform(0, 3, 200)], axis=1)
    f = np.zeros((200, 6), np.complex128)
    for k, (E0, w, cyc, t) in enumerate(args):
        p = Pulse()
        R.beyondDipolePulse_init(C.byref(p), E0, w, cyc)
        R.beoyndDipolePulse_axialPart(C.byref(p), t, f[k])
    out = os.path.join(ROOT, "tests", "golden", "reference_pulse_ik.npz")
    np.savez_compressed(out, index=idx, ik=ik, imu=imu, pulse_args=args, pulse_f=f)
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
