"""Build the engine library for gfx950 with hipcc (in-tree, no JIT cache).

    python -m opendust_amd.build [--save-temps DIR]

produces ``opendust_amd/lib/libopendust_amd.so``.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRCS = [os.path.join(HERE, "csrc", "od_engine.hip"), os.path.join(HERE, "csrc", "od_output.cpp")]
INCLUDE = os.path.join(ROOT, "include")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libopendust_amd.so")
ARCH = "gfx950"


def hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def compile_flags():
    return [
        "--offload-arch=" + ARCH,
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        # exact IEEE ordering everywhere; the pair kernels use explicit fma
        "-ffp-contract=off",
        # never emit packed FP32 VALU ops (v_pk_fma_f32 and friends)
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", INCLUDE,
        "-I", os.path.join(HERE, "csrc"),
    ]


def build(verbose=False, save_temps=None):
    """Compile the engine; always recompiles (the 'does it build' check)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = [hipcc()] + compile_flags() + SRCS + ["-o", tmp]
    if save_temps:
        os.makedirs(save_temps, exist_ok=True)
        cmd.insert(1, "-save-temps=obj")
        cmd[cmd.index("-o") + 1] = os.path.join(save_temps, "libopendust_amd.so")
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         cwd=save_temps or None)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed (%d):\n%s" % (res.returncode, res.stdout))
    if save_temps:
        return os.path.join(save_temps, "libopendust_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    st = None
    if "--save-temps" in sys.argv:
        st = os.path.abspath(sys.argv[sys.argv.index("--save-temps") + 1])
    print(build(verbose=True, save_temps=st))
