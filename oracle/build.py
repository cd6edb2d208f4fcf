"""Builds the oracle's C restatement (field_ref.c) into oracle/_build/libfieldref.so.

The reference itself is not built: its solver needs tira, Eigen, MKL, FFTW, Boost and CUDA, none of
which exists in this image, so no oracle/_ref is produced.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "libfieldref.so")


def build() -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    subprocess.run(["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-shared", "-fPIC", "-std=c11",
                    os.path.join(HERE, "field_ref.c"), "-o", tmp, "-lm"], check=True)
    os.replace(tmp, OUT)
    return OUT


def load():
    """ctypes handle of the built library (builds it first when missing)."""
    import ctypes as C
    if not os.path.exists(OUT):
        build()
    L = C.CDLL(OUT)
    P, I, F = C.c_void_p, C.c_int, C.c_float
    for name in ("cw_eval_grid_f32", "cw_eval_grid_f64"):
        fn = getattr(L, name)
        fn.restype = None
        fn.argtypes = [P, P, P, P, I, I, F, F, F, F, F, F, I, I, I, P]
    return L


if __name__ == "__main__":
    print(build())
