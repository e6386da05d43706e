"""Write the golden vectors under tests/golden/ from the genuine reference.

Runs only where the reference checkout exists (it reads its data files and uses
oracle/_ref/libckfft_ref.so, built from the reference's own sources by
``make -C oracle ref``).  The outputs are data: inputs and the reference's outputs.

  input_txt.f32      src/test/input.txt (4096 complex values, re/im interleaved),
                     the reference's own regression input
  example_in.f32     the 1024-float input array of src/example/main.cpp:39-42
  golden_small.npz   full reference outputs on those inputs for every n = 1..4096:
                     cfwd_<n>, cinv_<n>   complex forward/inverse of input_txt[:n]
                                          (maxCount = n; maxCount = 8192 checked equal)
                     rfwd_<n>             real forward of Re(input_txt[:n])
                     rinv_<n>             real inverse of rfwd_<n>
                     ex_rfwd, ex_rinv     the example's forward and round trip (n = 1024)
  golden_digests.json
                     SHA-256 of the reference's output bytes for n = 8192..65536 on
                     seeded inputs (oracle.uniform_pm1 / oracle.normal01), plus 64
                     sampled elements each, and SHA-256 of the twiddle tables the
                     reference's context builds for maxCount = 2^k, k <= 24.

Twiddle bits depend on the platform's cosf/sinf; these were produced with the libm
recorded in the json ("libm").
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import platform
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

REF_ROOT = os.environ.get("CKFFT_REFERENCE", "/root/reference")


def read_input_txt() -> np.ndarray:
    with open(os.path.join(REF_ROOT, "src", "test", "input.txt")) as f:
        vals = [float(line) for line in f if line.strip()]
    return np.array(vals, dtype=np.float32)


def read_example_input() -> np.ndarray:
    with open(os.path.join(REF_ROOT, "src", "example", "main.cpp")) as f:
        text = f.read()
    body = text[text.index("float input[] ="):]
    body = body[body.index("{") + 1: body.index("}")]
    vals = [float(t.rstrip("f")) for t in re.findall(r"[-+]?[0-9.]+(?:e[-+]?[0-9]+)?f", body)]
    assert len(vals) == 1024, len(vals)
    return np.array(vals, dtype=np.float32)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def samples(a: np.ndarray, k: int = 64):
    flat = a.ravel()
    idx = np.unique(np.linspace(0, flat.size - 1, k).astype(np.int64))
    return {"idx": idx.tolist(), "bits": flat[idx].view(np.uint32).tolist()}


def main():
    R = O.Reference()
    inp = read_input_txt()
    ex = read_example_input()
    inp.tofile(os.path.join(HERE, "input_txt.f32"))
    ex.tofile(os.path.join(HERE, "example_in.f32"))

    small = {}
    for k in range(0, 13):
        n = 1 << k
        x = inp[: 2 * n]
        ctx = R.init(n, 3)
        ctx_big = R.init(8192, 3)
        for inv, tag in ((False, "cfwd"), (True, "cinv")):
            y = R.complex(ctx, x, inv)
            y2 = R.complex(ctx_big, x, inv)
            assert np.array_equal(y.view(np.uint32), y2.view(np.uint32)), (tag, n)
            small[f"{tag}_{n}"] = y
        xr = np.ascontiguousarray(x[0::2])
        yr = R.real_forward(ctx, xr[None])[0]
        assert np.array_equal(yr.view(np.uint32), R.real_forward(ctx_big, xr[None])[0].view(np.uint32))
        small[f"rfwd_{n}"] = yr
        small[f"rinv_{n}"] = R.real_inverse(ctx, yr[None], n)[0]
        R.shutdown(ctx)
        R.shutdown(ctx_big)
    ctx = R.init(1024, 3)
    small["ex_rfwd"] = R.real_forward(ctx, ex[None])[0]
    small["ex_rinv"] = R.real_inverse(ctx, small["ex_rfwd"][None], 1024)[0]
    R.shutdown(ctx)
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **small)

    dig = {"libm": platform.libc_ver(), "machine": platform.machine(), "outputs": {}, "twiddles": {}}
    for k in range(13, 17):
        n = 1 << k
        ctx = R.init(n, 3)
        xc = O.uniform_pm1(0xC0 + k, 2 * n)
        xr = O.normal01(0xD0 + k, n)
        for inv, tag in ((False, "cfwd"), (True, "cinv")):
            y = R.complex(ctx, xc, inv)
            dig["outputs"][f"{tag}_{n}"] = {"seed": 0xC0 + k, "input": "uniform_pm1",
                                            "sha256": sha(y), **samples(y)}
        yr = R.real_forward(ctx, xr[None])[0]
        dig["outputs"][f"rfwd_{n}"] = {"seed": 0xD0 + k, "input": "normal01", "sha256": sha(yr), **samples(yr)}
        yi = R.real_inverse(ctx, yr[None], n)[0]
        dig["outputs"][f"rinv_{n}"] = {"seed": 0xD0 + k, "input": "rfwd of normal01",
                                       "sha256": sha(yi), **samples(yi)}
        R.shutdown(ctx)

    # Twiddle tables: the reference keeps them inside its context; read them back
    # through the context layout (tables follow an 8-byte-aligned header,
    # src/ckfft/context.cpp:24-114) via the user-buffer mode.
    lib = R.lib
    for k in range(0, 25):
        m = 1 << k
        size = ctypes.c_size_t(0)
        assert not lib.CkFftInit(m, 3, None, ctypes.byref(size))
        buf = np.zeros(size.value, np.uint8)
        ctx = lib.CkFftInit(m, 3, buf.ctypes.data, ctypes.byref(size))
        assert ctx == buf.ctypes.data
        header = size.value - 2 * 8 * m
        fwd = buf[header: header + 8 * m].view(np.float32)
        inv = buf[header + 8 * m:].view(np.float32)
        dig["twiddles"][str(m)] = {"header_bytes": int(header), "fwd_sha256": sha(fwd), "inv_sha256": sha(inv),
                                   **samples(fwd)}
    with open(os.path.join(HERE, "golden_digests.json"), "w") as f:
        json.dump(dig, f, indent=1)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
