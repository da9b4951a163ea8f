"""Build libbmx_hip.so for gfx950 in place (hipcc via biomedisa_amd/csrc/Makefile)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8, force=True):
    csrc = os.path.join(HERE, "csrc")
    if force:
        subprocess.run(["make", "-s", "-C", csrc, "clean"], check=True)
    subprocess.run(["make", "-s", "-j%d" % jobs, "-C", csrc], check=True)
    return os.path.join(HERE, "libbmx_hip.so")


if __name__ == "__main__":
    print(build())
