"""Where the on-chip encode kernel (k_encode_b32) spends its time: phase
clocks of a profiling build of the library.

    python -m idx2_amd.phase_profile --build    # on the build host (hipcc)
    python -m idx2_amd.phase_profile            # on the GPU: prints a markdown table

The profiling build (-DIDX2G_PHASE_TIMERS, idx2_amd/_build_phase) has thread 0
of every workgroup read the 100 MHz wall clock right after each workgroup
barrier that ends a phase and add the ticks since its previous mark to the
phase's counter; the sums over all workgroups, divided by the number of
workgroups, are the time one CU spends in each phase.  The workload is the
1024^3 float32 field (brick 32^3, 3 levels, tolerance 1e-4); one untimed
encode first, then one measured.  The marks add a few global atomics per
brick, so the profiling build runs slightly slower than the product.
"""
import argparse
import ctypes
import os
import sys

PHASES = [
    (0, "gather + min/max"), (1, "extrapolate: X (extended)"), (2, "extrapolate: Y (extended)"),
    (3, "extrapolate: Z (extended)"), (4, "extrapolate: levels 1-4 (17^3 .. 3^3 subgrids)"),
    (5, "extrapolate: Z inverse"), (6, "extrapolate: Y inverse"), (7, "extrapolate: X inverse"),
    (8, "forward: X"), (9, "forward: Y"), (10, "forward: Z"), (11, "parent copy + subband weights"),
    (12, "code: stage subband into LDS"), (13, "code: block stage"), (14, "code: stream layout + emission"),
    (15, "code: copy into the arena + loop tail"),
]
DEC_PHASES = [
    (16, "decode: setup (exponents, plane counts, parent rows issued)"),
    (17, "decode: z-odd subbands, inverse block stage into the parking area"),
    (18, "decode: z-odd subbands, wait for every wave"),
    (19, "decode: z-even subbands, inverse block stage"), (20, "decode: z-even subbands into the columns"),
    (21, "decode: unpark + inverse weights"), (22, "decode: inverse CDF5/3 (Z, Y, X)"),
    (23, "decode: crop into the output"),
]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--build", action="store_true", help="build the profiling library and exit")
    ap.add_argument("--dims", type=int, nargs=3, default=(1024, 1024, 1024))
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from idx2_amd import build as B
    if args.build:
        print(B.build(phase_timers=True))
        return
    import torch
    from idx2_amd import api
    if not os.path.exists(B.PHASE_LIB):
        raise SystemExit(f"{B.PHASE_LIB} is missing: run with --build on the build host first")
    api.LIB_PATH = B.PHASE_LIB
    L = api.lib()
    L.idx2g_phase_times.argtypes = [ctypes.c_void_p]
    ctx = api.Context(0)
    dims = tuple(args.dims)
    x = torch.empty(dims[0] * dims[1] * dims[2], dtype=torch.float32, device="cuda")
    ctx.synth_field(x.data_ptr(), dims, "float32", 0)
    torch.cuda.synchronize()
    ticks = (ctypes.c_ulonglong * 32)()
    for _ in range(2):
        L.idx2g_phase_times(ticks)  # clears
        ds = api.Dataset.encode(ctx, dev_ptr=x.data_ptr(), dims=dims, dtype="float32", brick=(32, 32, 32),
                                nlevels=3, tolerance=1e-4)
        torch.cuda.synchronize()
        st = ctx.stats()
        ds.free()
    L.idx2g_phase_times(ticks)
    nwg = torch.cuda.get_device_properties(0).multi_processor_count
    # one decode of the whole grid: the on-chip decode kernel's phases
    out = torch.empty_like(x)
    ds = api.Dataset.encode(ctx, dev_ptr=x.data_ptr(), dims=dims, dtype="float32", brick=(32, 32, 32),
                            nlevels=3, tolerance=1e-4)
    ds.decode(out_dev_ptr=out.data_ptr())
    torch.cuda.synchronize()
    dticks = (ctypes.c_ulonglong * 32)()
    L.idx2g_phase_times(dticks)
    ds.decode(out_dev_ptr=out.data_ptr())
    torch.cuda.synchronize()
    dst = ctx.stats()
    L.idx2g_phase_times(dticks)
    ds.free()
    tot = sum(ticks[i] for i, _ in PHASES)
    print(f"workload: {dims[0]}x{dims[1]}x{dims[2]} float32, brick 32^3, 3 levels, tolerance 1e-4")
    print(f"encode total {st['total_ms']:.3f} ms, level-0 k_encode_b32 {st['dominant_ms']:.3f} ms "
          f"(profiling build); phases summed over the 3 levels' launches, per workgroup ({nwg} workgroups)")
    print()
    print("| phase | ms per CU | % |")
    print("|---|---:|---:|")
    for i, name in PHASES:
        print(f"| {name} | {ticks[i] * 10 / nwg / 1e6:.3f} | {100.0 * ticks[i] / max(tot, 1):.1f} |")
    print(f"| total | {tot * 10 / nwg / 1e6:.3f} | 100.0 |")
    dtot = sum(dticks[i] for i, _ in DEC_PHASES)
    print()
    print(f"decode total {dst['total_ms']:.3f} ms (profiling build); phases of the on-chip decode kernel "
          f"(k_decode_b32) summed over the 3 levels' launches, per workgroup")
    print()
    print("| phase | ms per CU | % |")
    print("|---|---:|---:|")
    for i, name in DEC_PHASES:
        print(f"| {name} | {dticks[i] * 10 / nwg / 1e6:.3f} | {100.0 * dticks[i] / max(dtot, 1):.1f} |")
    print(f"| total | {dtot * 10 / nwg / 1e6:.3f} | 100.0 |")


if __name__ == "__main__":
    main()
