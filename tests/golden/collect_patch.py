"""Copy the reference's box patch test (tension with prescribed displacement, one-partition
inputs of the linear- and quadratic-tet boxes) and its expected output into
tests/golden/patch_tension/ as test data: box_LT_0.in, box_LT_0.in.st, box_QT_0.in,
box_QT_0.in.st and stress_tension_disp.out (the domain average of the PK2 S_xx that
verification_tests/src/tension_disp.cc writes).  Data files of the reference's own test
(tests/box/tension_disp, args -SS -disp -V), copied byte for byte."""
import os
import shutil
import sys

SRC = "/root/reference/verification_tests"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "patch_tension")


def main():
    if not os.path.isdir(SRC):
        sys.exit(f"{SRC} not found")
    os.makedirs(DST, exist_ok=True)
    for lt in ("LT", "QT"):
        d = os.path.join(SRC, f"box_patch_{lt}", "tension_disp", "inputs", f"box_{lt}_1CPU")
        for fn in (f"box_{lt}_0.in", f"box_{lt}_0.in.st"):
            shutil.copyfile(os.path.join(d, fn), os.path.join(DST, fn))
    shutil.copyfile(os.path.join(SRC, "output", "stress_tension_disp.out"), os.path.join(DST, "stress_tension_disp.out"))


if __name__ == "__main__":
    main()
