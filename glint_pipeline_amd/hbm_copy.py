"""One device-to-device copy of a known size through libglint's streaming copy
kernel (k_copy).  Profiled with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE it
gives the counter reading of a plain copy, which scales the same counters
read on the contact-tracing kernels to bytes.

    python -m glint_pipeline_amd.hbm_copy --bytes 1073741824
"""
from __future__ import annotations

import argparse

from . import runtime


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--bytes", type=int, default=1 << 30)
    ap.add_argument("--repeat", type=int, default=1)
    a = ap.parse_args()
    src = runtime.DeviceBuffer(a.bytes)
    dst = runtime.DeviceBuffer(a.bytes)
    for _ in range(a.repeat):
        st = runtime.lib().glint_device_copy(dst.ptr, src.ptr, a.bytes)
        if st:
            raise SystemExit(f"glint_device_copy failed ({st})")
    print(f"copied {a.bytes} bytes x {a.repeat}")


if __name__ == "__main__":
    main()
