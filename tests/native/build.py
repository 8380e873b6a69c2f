"""Builds tests/native/libf32math_probe.so (test harness for the f32 math
routines of the shading code) with the library's own hipcc flags."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "f32math_probe.hip")
LIB = os.path.join(HERE, "libf32math_probe.so")


def build():
    from voidray_amd import build as vb
    tmp = LIB + ".tmp"
    cmd = [vb._hipcc(), "-x", "hip", *vb.HIP_FLAGS, "-shared", SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"probe build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
