"""Generate the pass-kernel instantiation units and their registries.

Each pass shape (a run of radix stages, stage 0 = unit sub-stride) is
instantiated for the four pass kinds (A^T, A, conj(A), fused convolution)
and both tile layouts (strided lines / contiguous lines), in FP64 and FP32.
The registry rows let the C host pick a kernel by (precision, kind, layout,
stage list).  Run by the package build; output goes to csrc/gen/.
"""
from __future__ import annotations

import os
import sys

KINDS = ["PACG_K_DIF", "PACG_K_DIT", "PACG_K_DITC", "PACG_K_FUSED"]
LAYOUTS = ["PACG_L_STRIDED", "PACG_L_CONTIG"]

# pure-radix runs (every pass of a pure radix-r schedule) ...
SHAPES = [[2] * k for k in range(1, 9)] + [[4] * k for k in range(1, 5)] + [[8] * k for k in range(1, 4)]
# ... and the mixed runs of the mixed-radix schedules the planner uses:
# 256 = 4*8*8 and 8*8*4, 2048 = 4^5 * 2 split as (4^4)(4, 2), 32 = 4*8 / 8*4
SHAPES += [[4, 8, 8], [8, 8, 4], [4, 2], [2, 4], [4, 8], [8, 4], [8, 2], [2, 8]]
# whole 2048-point axes in one pass on contiguous lines: 4^5 * 2, 2^11 (the
# reference's radix for 2048) and 8^3 * 4 / 4 * 8^3
LONG_SHAPES = [[4, 4, 4, 4, 4, 2], [2] * 11, [8, 8, 8, 4], [4, 8, 8, 8]]
# ... and on strided lines in FP64 (the strided axis of a 2048^2 array in one
# pass: 3 round trips per 2D convolution instead of 5).  A tile is 2, 4 or 8
# lines of 2048 points (32 KiB each in FP64), so few lines fit in LDS: points
# per thread -> (lines per tile, rank) compiled, rank 0 the default.  16 points
# on 8 lines (64-byte runs of each row) are fastest at 2048^2 (about 37 against
# 50 us with 8 points on 2 lines); no fused pass with 16 points per thread
# (data and spectrum would spill), so the fused pass defaults to 8 on 2 lines.
LONG_STRIDED = [[4, 4, 4, 4, 4, 2]]
LONG_STRIDED_TILES = {16: [(8, 0), (4, 1)], 8: [(2, 2), (4, 3)]}
LONG_STRIDED_FUSED = [(2, 0), (4, 1)]

E_TARGET = 16


def phases(shape, et=E_TARGET):
    """Greedy bottom-up grouping of stages into register phases (product <= et)."""
    out = []
    cur = 1
    for g, r in enumerate(shape):
        if cur > 1 and cur * r > et:
            out.append(g - 1)
            cur = 1
        cur *= r
    out.append(len(shape) - 1)
    return out


def codes(shape, et=E_TARGET):
    lg = {2: 1, 4: 2, 8: 3}
    sc = 0
    for g, r in enumerate(shape):
        sc |= lg[r] << (2 * g)
    ph = phases(shape, et)
    pc = 0
    for p, top in enumerate(ph):
        pc |= top << (4 * p)
    return sc, len(ph), pc


def geometry(shape, prec, et=E_TARGET):
    P = 1
    for r in shape:
        P *= r
    ph = phases(shape, et)
    E = 1
    lo = 0
    for top in ph:
        R = 1
        for r in shape[lo:top + 1]:
            R *= r
        E = max(E, R)
        lo = top + 1
    tpl = P // E
    tj = 32 if prec == "float" else 16
    strided_lines = max(tj, 256 // tpl) if tpl < 8 else tj
    while strided_lines * tpl > 1024:
        strided_lines //= 2
    elem = 4 if prec == "float" else 8
    while P * (strided_lines + 1) * 2 * elem > 80 * 1024 and strided_lines > 4:
        strided_lines //= 2
    nl = max(1, 256 // tpl)
    ldp = P + P // 8 + 1
    while nl > 1 and nl * ldp * 2 * elem > 40 * 1024:
        nl //= 2
    return P, E, tpl, strided_lines, nl


def tile_options(shape, prec, layout, et=E_TARGET):
    """Lines per workgroup to compile: the default first, then the autotuner's alternatives."""
    P, E, tpl, sl, nl = geometry(shape, prec, et)
    elem = 4 if prec == "float" else 8
    if layout == "PACG_L_STRIDED":
        opts = [sl]
        if P >= 64:
            for alt in (sl // 2, sl // 4):
                if alt >= 4 and alt * tpl >= 64:
                    opts.append(alt)
            # lines of up to 256 points: narrower tiles, so that small arrays
            # (a 2^16-point line has 256 lines per pass) still fill the CUs
            if P <= 256:
                for alt in (sl // 8, sl // 16):
                    if 1 <= alt < 4 and alt * tpl >= 64:
                        opts.append(alt)
    else:
        opts = [nl]
        if P >= 64:
            alts = (nl // 2, nl * 2) + ((nl // 4,) if P <= 256 else ())
            for alt in alts:
                if 1 <= alt and alt * tpl <= 1024 and alt * tpl >= 64 and alt * (P + P // 8 + 1) * 2 * elem <= 80 * 1024:
                    opts.append(alt)
    return opts


NCHUNK = 8


def emit(prec: str, outdir: str):
    """Write NCHUNK translation units for one precision; returns their table names."""
    T = prec
    tag = "f64" if prec == "double" else "f32"
    pconst = "PACG_F64" if prec == "double" else "PACG_F32"
    spec_types = ["double"] if prec == "double" else ["float", "double"]
    rows = []
    # FP64 long lines default to 8-point phases: more phases, but half the
    # registers per thread (7 instead of 4 waves per SIMD; 2048^2 passes 37.6
    # -> 33 us); the 16-point phases stay as autotuner alternatives
    e8 = [s for s in LONG_SHAPES if prec == "double" and phases(s, 8) != phases(s)]
    jobs = [(shape, E_TARGET, "", 8 if shape in e8 else 0) for shape in SHAPES + LONG_SHAPES]
    jobs += [(shape, 8, "_e8", 0) for shape in e8]
    for shape, et, suffix, rank0 in jobs:
        sc, npph, pc = codes(shape, et)
        P, E, tpl, sl, nl = geometry(shape, prec, et)
        shname = "Shape<%d, 0x%Xu, %d, 0x%Xu>" % (len(shape), sc, npph, pc)
        sname = "x".join(str(r) for r in shape) + suffix
        long_s = prec == "double" and shape in LONG_STRIDED
        for kind in KINDS:
            for layout in (LAYOUTS if shape in SHAPES or long_s else ["PACG_L_CONTIG"]):
                specs = spec_types if kind == "PACG_K_FUSED" else [T]
                if shape in SHAPES or layout == "PACG_L_CONTIG":
                    tiles = [(ln, rank0 + i) for i, ln in enumerate(tile_options(shape, prec, layout, et))]
                elif kind == "PACG_K_FUSED":
                    if E > 8:
                        continue
                    tiles = LONG_STRIDED_FUSED
                else:
                    tiles = LONG_STRIDED_TILES[E]
                for ln, rank in tiles:
                    for ts in specs:
                        fn = "launch_pass<%s, %s, %s, %s, %s, %d>" % (T, ts, shname, kind, layout, ln)
                        sp = "PACG_F64" if ts == "double" else "PACG_F32"
                        nm = "%s_%s_%s_%s_%s_l%d" % (tag, kind[7:].lower(), "s" if layout == "PACG_L_STRIDED" else "c",
                                                     sname, "h64" if (kind == "PACG_K_FUSED" and ts == "double") else "h",
                                                     ln)
                        rows.append('  {%s, %s, %s, %s, %d, 0x%Xu, 0x%Xu, %d, %d, %d, %d, &%s, "%s"},'
                                    % (pconst, sp, kind, layout, len(shape), sc, pc, P, ln, ln * tpl, rank, fn, nm))
    names = []
    for k in range(NCHUNK):
        part = rows[k::NCHUNK]
        nm = "pacg_variants_%s_%d" % (tag, k)
        names.append((nm, len(part)))
        lines = [
            "// generated by gen_variants.py -- pass kernel instantiations (%s, part %d)" % (tag, k),
            '#include "../pacg_pass.hpp"',
            "using namespace pacg;",
            'extern "C" {',
            "extern pacg_variant_t %s[];" % nm,
            "pacg_variant_t %s[] = {" % nm,
        ] + part + ["};", "}"]
        with open(os.path.join(outdir, "%s.hip" % nm), "w") as f:
            f.write("\n".join(lines) + "\n")
    return names


def main(outdir: str):
    os.makedirs(outdir, exist_ok=True)
    tables = emit("double", outdir) + emit("float", outdir)
    idx = ["/* generated by gen_variants.py -- registry of all pass kernels */",
           '#include "../pacg_internal.h"']
    for nm, _ in tables:
        idx.append("extern pacg_variant_t %s[];" % nm)
    idx.append("const pacg_variant_t* const pacg_variant_tables[] = {%s};" % ", ".join(nm for nm, _ in tables))
    idx.append("const int pacg_variant_counts[] = {%s};" % ", ".join(str(c) for _, c in tables))
    idx.append("const int pacg_num_variant_tables = %d;" % len(tables))
    with open(os.path.join(outdir, "pacg_variants_index.c"), "w") as f:
        f.write("\n".join(idx) + "\n")
    print("generated %d pass variants in %d units" % (sum(c for _, c in tables), len(tables)))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)), "gen"))
