"""Copy the reference's crystal-plasticity verification deck (tension case) and its restart
fixtures into tests/golden/cp_box_tension/ as test data: the 8-partition mesh inputs
(box_*.in), the solver file, the model parameters, the per-partition crystal orientations
(CO/co_*.in) and the step-49 restart files (ref_out_tension/*.res).  These are data files of
the reference's own test (verification_tests/crystal_plasticity), copied byte for byte."""
import os
import shutil
import sys

SRC = "/root/reference/verification_tests/crystal_plasticity"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cp_box_tension")


def main():
    if not os.path.isdir(SRC):
        sys.exit(f"{SRC} not found")
    os.makedirs(os.path.join(DST, "CO"), exist_ok=True)
    os.makedirs(os.path.join(DST, "ref_out"), exist_ok=True)
    for p in range(8):
        shutil.copyfile(os.path.join(SRC, "box_8CPU_tension", f"box_{p}.in"), os.path.join(DST, f"box_{p}.in"))
        shutil.copyfile(os.path.join(SRC, "CO", f"co_{p}.in"), os.path.join(DST, "CO", f"co_{p}.in"))
        shutil.copyfile(os.path.join(SRC, "ref_out_tension", f"box_{p}_49.res"),
                        os.path.join(DST, "ref_out", f"box_{p}_49.res"))
    shutil.copyfile(os.path.join(SRC, "box_0.in.st_tension"), os.path.join(DST, "box.in.st"))
    shutil.copyfile(os.path.join(SRC, "box_8CPU_tension", "model_params.in"), os.path.join(DST, "model_params.in"))
    shutil.copyfile(os.path.join(SRC, "ref_out_tension", "time_step_info.res"),
                    os.path.join(DST, "ref_out", "time_step_info.res"))


if __name__ == "__main__":
    main()
