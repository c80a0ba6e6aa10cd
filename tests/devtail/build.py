"""Build tests/devtail/libdevtail.so (gfx950) with the library's compiler flags."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libdevtail.so")


def build(verbose: bool = False) -> str:
    from dbgenzpaper_amd import build as B

    cmd = [B.HIPCC, *B.CFLAGS, "-shared", "-I", B.CSRC, os.path.join(HERE, "devtail.hip"), "-o", LIB]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for devtail.hip:\n" + r.stderr[-4000:])
    if verbose:
        print("built", LIB)
    return LIB
