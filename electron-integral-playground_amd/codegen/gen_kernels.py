"""Emit straight-line HIP kernels, one per angular class, for the McMurchie-Davidson path.

In the style of the reference's offline generator (gen/equation_mcmurchie_davidson.py:4-98), which
turns the MD recursions into straight-line C, but targeting gfx950 and restructured for it:

* ``pairE_<la><lb>``: one thread per primitive pair.  Builds the pair's Hermite expansion E^{ij}_t by
  the two-index Hermite recurrence (the coefficients of reference mcmurchie_davidson_E_term.hpp:22-45,
  166-199, without its transfer between centres), contracts it straight into the *output* basis of the
  pair (real solid harmonics, or normalised Cartesians) and folds the pair coefficient in.  Entries are
  stored t-major, in exactly the order the three-centre kernels consume them, so those kernels read
  them as wave-uniform scalar loads.
* ``j3c_<la><lb><lc>_<m0>_<m1>``: one wave per contracted shell pair, one lane per auxiliary shell.
  Each lane loops over the shell pair's primitive pairs (uniform) and its own auxiliary primitives,
  evaluates the Boys function and the Hermite integrals R_tuv by the MD recursion in registers, applies
  the auxiliary shell's Hermite-to-output transform (for spherical output only the |s| = lc terms
  survive -- Hobson's theorem -- so it is a small constant matrix) and accumulates
  out[ij, m] += E[ij, t] * V[t, m] in registers.  Components m0..m1 of the auxiliary shell are handled
  by one kernel, so large classes are split over several kernels to bound register use.  The result is
  identical to the reference's per-primitive Cartesian evaluation followed by normalisation and the
  spherical transform (three_center.cpp:27-289) up to rounding.
* ``v_<la><lb>``: the same structure with point charges in place of auxiliary shells (lc = 0,
  zeta = p or p w^2/(p + w^2)), reference nuclear_attraction.cpp:23-171.
* ``j2c_<li><lj>``: one lane per auxiliary shell pair; both sides single-centre
  (reference two_center.cpp:25-212).
"""


from .tables import (aux_transform, cart_list, fmt, herm_list, n_out, ncart, pair_out_rows,
                     pair_support, SqrtRational)


# ---------------------------------------------------------------------------------------------
# Layout of the packed pair matrix
# ---------------------------------------------------------------------------------------------

def pair_layout(la, lb, spherical):
    """Entries (t, row) of the packed pair matrix in storage order (t-major)."""
    sup = pair_support(la, lb, spherical)
    ts = sorted({t for s in sup for t in s}, key=lambda t: (sum(t), -t[0], -t[1]))
    entries = []
    for t in ts:
        for row, s in enumerate(sup):
            if t in s:
                entries.append((t, row))
    return entries


def pair_nnz(la, lb, spherical):
    return len(pair_layout(la, lb, spherical))


def pair_stride(la, lb, spherical):
    """Row stride of the pair matrix: the entries rounded up to 4, so every row starts on a 16-byte
    boundary in both the float64 and float32 copies (rows are staged into LDS in 16-byte pieces)."""
    return (pair_nnz(la, lb, spherical) + 3) // 4 * 4


# Three-centre kernels stage each primitive pair's matrix row (and P) into LDS with asynchronous
# global->LDS loads, double-buffered: the row of the next primitive pair is in flight while the current
# one is contracted, and the contraction reads wave-uniform LDS broadcasts instead of waiting on scalar
# loads.
STAGE_LDS = True
STAGE_BATCH_BYTES = 2048      # rows of small classes are staged this many bytes at a time ...
STAGE_MAX_ROWS = 16           # ... but at most this many rows (P of each row: 32 bytes, 2 lanes)

# How a three-centre kernel feeds the wave-uniform pair-matrix entries to the contraction:
# "lds": rows staged into LDS as above, each entry read back as an LDS broadcast;
# "dpp": the row of a primitive pair lives in registers spread over the lanes -- entry k in VGPR k / 16
#        at lane k % 16 of every row of 16 lanes (the 4 rows hold the same 16 entries) -- loaded with
#        plain vector loads one primitive pair ahead, and each entry reaches all 64 lanes by a DPP row
#        broadcast (v_mov_b64_dpp row_newbcast), so the contraction never waits on memory.
STAGE = "lds"
# Order of the Hermite integrals R^0_t in the three-centre kernels: "level" -- the whole MD recursion,
# level by level, before the contraction; "demand" -- each R^0_t (with what it depends on) emitted just
# before the first auxiliary-transform term that reads it, so that values die sooner.  Measured on the
# 200-atom slice (float64): 48 chunk kernels gained a wave of occupancy (e.g. (dd|f), (pd|f) chunks 1 -> 2,
# 20-30 % fewer registers), none lost one; classes timed alone 4.6 % faster in sum ((df|g) 96.6 -> 86.9
# ms, (pd|f) 32.5 -> 30.2 ms).
R_ORDER = "demand"
# Rows longer than this many entries stay in LDS even in "dpp" mode (the register copy of a row costs
# two VGPRs per 16 entries).
DPP_MAX_ENTRIES = 512
# Float64 "dpp" kernels whose chunk keeps at most this many doubles live (accumulators + Hermite
# integrals) can pipeline the per-triplet prologue: the next primitive pair's zeta, prefactor, T and
# Boys grid row are computed and fetched one iteration ahead (its centre two ahead), so the table gather
# overlaps the current contraction.  Measured with a limit of 140 on the plain and 200-atom workloads:
# 5-49 % slower per class (the ~30 extra registers cost occupancy), so it is off (limit 0).
PREFETCH_MAX_DOUBLES = 0


# ---------------------------------------------------------------------------------------------
# Hermite E (pair) code: two-index recurrence per axis
# ---------------------------------------------------------------------------------------------

def _emit_E_axis(lines, ax, la, lb, ty="double"):
    """Emit E{ax}_{i}_{j}_{t} for i <= la, j <= lb, t <= i + j by the two-index Hermite recurrence

        E^{i+1,j}_t = 1/(2p) E^{ij}_{t-1} + PA E^{ij}_t + (t+1) E^{ij}_{t+1}   (first index), then
        E^{i,j+1}_t = 1/(2p) E^{ij}_{t-1} + PB E^{ij}_t + (t+1) E^{ij}_{t+1}   (second index).

    The reference builds E^{i+j,0} and transfers to the second centre, E^{ij} = sum_m C(j, m) AB^{j-m}
    E^{i+m,0} (mcmurchie_davidson_E_term.hpp:166-199): the same coefficients, but for a diffuse A and a
    tight B (P near B, |AB| of several Bohr) the transfer's terms of order AB^{i+j} cancel to a value
    smaller by many orders (seen as 1e-4 relative error on a (df|s) integral in float64).  The
    recurrence in j uses PB, which is small there, and cancels nothing."""
    _emit_E_axis_two_index(lines, ax, la, lb, ty)


def _emit_E_axis_two_index(lines, ax, la, lb, ty="double"):
    """E{ax}_{i}_{j}_{t} by the recurrences in i (with PA) and j (with PB); see _emit_E_axis."""
    def name(i, j, t):
        return f"E{ax}_{i}_{j}_{t}"

    def step(i, j, t, X, src):
        terms = []
        if t >= 1:
            terms.append(f"o2p * {name(*src, t - 1)}")
        if t <= sum(src):
            terms.append(f"{X}{ax} * {name(*src, t)}")
        if t + 1 <= sum(src):
            terms.append(f"{fmt(t + 1)} * {name(*src, t + 1)}")
        lines.append(f"    const {ty} {name(i, j, t)} = {' + '.join(terms)};")

    lines.append(f"    const {ty} {name(0, 0, 0)} = 1.0;")
    for i in range(la):
        for t in range(i + 2):
            step(i + 1, 0, t, "PA", (i, 0))
    for i in range(la + 1):
        for j in range(lb):
            for t in range(i + j + 2):
                step(i, j + 1, t, "PB", (i, j))


def emit_pairE(la, lb, spherical):
    name = f"pairE_{la}{lb}_{'s' if spherical else 'c'}"
    ca, cb = cart_list(la), cart_list(lb)
    rows = pair_out_rows(la, lb, spherical)
    layout = pair_layout(la, lb, spherical)
    nnz = len(layout)
    L = ["// pair matrix of class (%d,%d) %s: %d entries per primitive pair" %
         (la, lb, "spherical" if spherical else "cartesian", nnz)]
    L.append(f'extern "C" __global__ void __launch_bounds__(256) {name}(const PairPrepArgs a)')
    L.append("{")
    L.append("    const int pp = blockIdx.x * blockDim.x + threadIdx.x;")
    L.append("    if (pp >= a.n_pp) return;")
    L.append("    const double* __restrict__ Pp = a.P_p + 4 * (size_t) pp;")
    L.append("    const double* __restrict__ Aa = a.A_a + 4 * (size_t) pp;")
    L.append("    const double* __restrict__ Bb = a.B_b + 4 * (size_t) pp;")
    L.append("    const double c = a.coef[pp];")
    L.append("    double Px, Py, Pz, p;")
    L.append("    if (a.recompute_P) {   // overlap ignores the stored P (reference overlap.cpp:22-29)")
    L.append("        p = Aa[3] + Bb[3];")
    L.append("        const double f = -Bb[3] / p;")
    L.append("        Px = Aa[0] + f * (Aa[0] - Bb[0]); Py = Aa[1] + f * (Aa[1] - Bb[1]); Pz = Aa[2] + f * (Aa[2] - Bb[2]);")
    L.append("    } else { Px = Pp[0]; Py = Pp[1]; Pz = Pp[2]; p = Pp[3]; }")
    L.append("    const double o2p = 0.5 / p;")
    L.append("    const double PAx = Px - Aa[0], PAy = Py - Aa[1], PAz = Pz - Aa[2];")
    L.append("    const double PBx = Px - Bb[0], PBy = Py - Bb[1], PBz = Pz - Bb[2];")
    L.append("    (void) PAx; (void) PAy; (void) PAz; (void) PBx; (void) PBy; (void) PBz; (void) o2p;")
    for ax in "xyz":
        _emit_E_axis(L, ax, la, lb)
    stride = pair_stride(la, lb, spherical) if spherical else nnz
    L.append("    double* __restrict__ out = a.E + (size_t) pp * %d;" % stride)
    L.append("    float* __restrict__ out32 = a.E32 ? a.E32 + (size_t) pp * %d : nullptr;" % stride)
    for k, (t, row) in enumerate(layout):
        r, terms = rows[row]
        pieces = []
        for (ci, cj), q in terms.items():
            ix, iy, iz = ca[ci]
            jx, jy, jz = cb[cj]
            if t[0] > ix + jx or t[1] > iy + jy or t[2] > iz + jz:
                continue
            coef = SqrtRational(r, q).to_float()
            prod = f"Ex_{ix}_{jx}_{t[0]} * Ey_{iy}_{jy}_{t[1]} * Ez_{iz}_{jz}_{t[2]}"
            pieces.append(prod if coef == 1.0 else f"{fmt(coef)} * ({prod})")
        expr = " + ".join(pieces) if pieces else "0.0"
        L.append(f"    {{ const double v = c * ({expr}); out[{k}] = v; if (out32) out32[{k}] = (float) v; }}")
    L.append("}")
    return name, stride, "\n".join(L)


# ---------------------------------------------------------------------------------------------
# Hermite integrals R_tuv by the MD recursion (straight-line, only what is needed)
# ---------------------------------------------------------------------------------------------

def _r_name(n, t, prefix="r"):
    return f"{prefix}{n}_{t[0]}_{t[1]}_{t[2]}"


def _R_deps(needed0):
    """MD recursion graph of R^0_t for t in needed0: {(n, t): [(n + 1, t - e), (n + 1, t - 2e)]}."""
    need = set((0, t) for t in needed0)
    stack = list(need)
    deps = {}
    while stack:
        n, t = stack.pop()
        tx, ty_, tz = t
        if t == (0, 0, 0):
            deps[(n, t)] = []
            continue
        if tx > 0:
            d = [(n + 1, (tx - 1, ty_, tz))] + ([(n + 1, (tx - 2, ty_, tz))] if tx > 1 else [])
        elif ty_ > 0:
            d = [(n + 1, (tx, ty_ - 1, tz))] + ([(n + 1, (tx, ty_ - 2, tz))] if ty_ > 1 else [])
        else:
            d = [(n + 1, (tx, ty_, tz - 1))] + ([(n + 1, (tx, ty_, tz - 2))] if tz > 1 else [])
        deps[(n, t)] = d
        for x in d:
            if x not in need:
                need.add(x)
                stack.append(x)
    return deps


def _R_line(n, t, ty, base="R000_", prefix="r"):
    tx, ty_, tz = t
    if t == (0, 0, 0):
        return f"        const {ty} {_r_name(n, t, prefix)} = {base}{n};"
    if tx > 0:
        a, b, X, k = (tx - 1, ty_, tz), (tx - 2, ty_, tz), "PQx", tx - 1
    elif ty_ > 0:
        a, b, X, k = (tx, ty_ - 1, tz), (tx, ty_ - 2, tz), "PQy", ty_ - 1
    else:
        a, b, X, k = (tx, ty_, tz - 1), (tx, ty_, tz - 2), "PQz", tz - 1
    expr = f"{X} * {_r_name(n + 1, a, prefix)}"
    if k > 0:
        expr += f" + {fmt(k)}{'f' if ty == 'float' else ''} * {_r_name(n + 1, b, prefix)}"
    return f"        const {ty} {_r_name(n, t, prefix)} = {expr};"


class _RDemand:
    """Emits R^0_t (and what it depends on) when first requested, depth first, so that the Hermite
    integrals are computed next to the contraction that consumes them."""

    def __init__(self, needed0, ty):
        self.deps, self.ty, self.done = _R_deps(needed0), ty, set()

    def request(self, lines, t):
        stack = [((0, t), False)]
        while stack:
            node, expanded = stack.pop()
            if node in self.done:
                continue
            if expanded:
                self.done.add(node)
                lines.append(_R_line(node[0], node[1], self.ty))
                continue
            stack.append((node, True))
            for d in reversed(self.deps[node]):
                if d not in self.done:
                    stack.append((d, False))


def _emit_R(lines, L, needed0, ty, base="R000_", prefix="r"):
    """Emit R^0_t for t in needed0 (plus dependencies), from R^n_000 named <base>n, level by level."""
    for n, t in sorted(_R_deps(needed0), key=lambda x: (-x[0], sum(x[1]), x[1])):
        lines.append(_R_line(n, t, ty, base, prefix))


# ---------------------------------------------------------------------------------------------
# Three-centre kernel
# ---------------------------------------------------------------------------------------------

def _prologue_lines(P, indent):
    """Per-triplet prologue of the primitive pair whose centre and exponent are P0..P3 (double): the
    separation from the auxiliary centre, zeta, the prefactor and T, into ndx/ndy/ndz, nzeta, npf, nT."""
    return [
        f"{indent}{{ const double dx = {P}0 - Cx, dy = {P}1 - Cy, dz = {P}2 - Cz, pe = {P}3;",
        f"{indent}  double z_, f_;",
        f"{indent}  if (a.omega == 0.0) {{ const double rs = eip::rsq_nr(pe + q); z_ = pe * q * (rs * rs); f_ = ckq * eip::rcp_nr(pe) * rs; }}",
        f"{indent}  else {{ const double w2 = a.omega * a.omega; z_ = pe * q * eip::rcp_nr(q + pe + pe * q * eip::rcp_nr(w2));",
        f"{indent}         f_ = ckq * a.omega * eip::rcp_nr(pe) * eip::rsq_nr(pe * q + (pe + q) * w2); }}",
        f"{indent}  nzeta = z_; npf = f_; ndx = dx; ndy = dy; ndz = dz; nT = z_ * (dx * dx + dy * dy + dz * dz); }}",
    ]


def _aux_prefactor_lines(lc, indent="      "):
    """Per auxiliary primitive (hoisted out of the primitive-pair loop): hl = (-1/(2q))^lc, the common
    factor of the auxiliary side's Hermite expansion, and ckq = ck 2 pi^(5/2) hl / q, the part of the
    Coulomb prefactor 2 pi^(5/2) / (p q sqrt(p + q)) that depends on the auxiliary primitive only."""
    lines = [f"{indent}const double inv_q = 1.0 / q;"]
    if lc > 0:
        lines.append(f"{indent}const double hq = -0.5 * inv_q;")
        lines.append(f"{indent}double hl = hq;")
        for _ in range(lc - 1):
            lines.append(f"{indent}hl *= hq;")
    else:
        lines.append(f"{indent}const double hl = 1.0;")
    lines.append(f"{indent}const double ckq = ck * eip::kTwoPi25 * hl * inv_q;")
    return lines



# Bytes of one order's coarse Boys slice staged in LDS (boys_table.py: 15 columns x 73 rows, padded)
BOYS_LDS_BYTES = 8768
BOYS_LDS_LMAX = 18


def boys_in_lds(la, lb, lc, precision, boys_lds):
    return bool(boys_lds) and precision != "f32" and la + lb + lc <= BOYS_LDS_LMAX


def emit_j3c(la, lb, lc, m0, m1, precision="f64", stage=None, boys_lds=False, r_order=None):
    """Spherical three-centre kernel for aux components [m0, m1) of class (la lb | lc); ``stage`` is
    "lds", "dpp", "ldsdpp" or "sgpr" (default STAGE), the way pair-matrix entries reach the lanes; ``boys_lds``:
    the float64 Boys function reads the class's coarse table slice staged in LDS (boys64_lds) instead of
    the fine grid in global memory."""
    spherical = True
    stage = stage or STAGE
    r_order = r_order or R_ORDER
    if stage not in ("lds", "dpp", "sgpr", "ldsdpp"):
        raise ValueError(f"unknown staging mode {stage!r}")
    # sgpr: no staging -- the wave-uniform entries are scalar loads used as SGPR operands of the FMAs;
    # ldsdpp: rows staged in LDS as for "lds", but read 16 entries per load (each lane of a row of 16 its
    # own entry) two groups ahead of use, and broadcast from those registers by DPP
    use_dpp = stage == "dpp" and pair_nnz(la, lb, spherical) <= DPP_MAX_ENTRIES
    use_lds = not use_dpp and STAGE_LDS and stage != "sgpr"
    lds_dpp = use_lds and stage == "ldsdpp"
    nij_ = n_out(la, spherical) * n_out(lb, spherical)
    prefetch = (use_dpp and precision != "f32"
                and nij_ * (m1 - m0) + _needed_R(la, lb, lc, m0, m1) <= PREFETCH_MAX_DOUBLES)
    f32 = precision == "f32"
    ty = "float" if f32 else "double"
    lit = (lambda v: fmt(v) + "f") if f32 else fmt
    name = f"j3c_{la}{lb}{lc}_{m0}_{m1}" + ("_f32" if f32 else "")
    L = la + lb + lc
    layout = pair_layout(la, lb, spherical)
    nnz = len(layout)
    nij = n_out(la, spherical) * n_out(lb, spherical)
    nj = n_out(lb, spherical)
    aux_rows = aux_transform(lc, spherical)
    ms = list(range(m0, m1))
    # Hermite t of the pair used by the layout, and R_{t+s} needed
    ts = []
    for t, _ in layout:
        if t not in ts:
            ts.append(t)
    needed = set()
    for t in ts:
        for m in ms:
            for s in aux_rows[m][1]:
                needed.add((t[0] + s[0], t[1] + s[1], t[2] + s[2]))
    out = [f"// ({la}{lb}|{lc}) aux components [{m0},{m1}) {precision}: L = {L}, {nnz} pair entries"]
    # One body, three callers: '{name}' reads the molecule's data from its kernel arguments (rematerialised
    # scalar loads from the kernarg segment), '{name}_jobs' from the per-item job table, and the
    # persistent queue kernels (one per register tier) call it once per work item they pull.  lb_ is the
    # item's index in the launch (items a.item_begin + lb_), smem the caller's LDS staging buffer
    # (j3c_smem_bytes(...) bytes, 16-byte aligned; unused by unstaged and DPP-staged bodies).
    # ArgsT: J3cArgs (the kernels' own by-value argument) or the constant-address-space J3cArgs of a queue
    # slot; either way the fields are invariant loads the compiler can re-issue where they are used.
    out.append("template <bool kJobs, class ArgsT>")
    # tid: the lane (threadIdx.x), passed in so a persistent caller can hand over a per-item copy (values
    # derived from it then stay inside the item instead of being hoisted out of the caller's loop).
    out.append(f"__device__ __forceinline__ void {name}_body(const ArgsT& a, const int lb_, char* __restrict__ smem, const int tid)")
    out.append("{")
    out.append("    const int item = a.item_begin + lb_;")
    out.append("    if (item >= a.item_end) return;")
    if not f32:
        out.append("    if (a.flags_in && !eip::item_flagged(a.flags_in + (size_t) item * a.flag_stride, a.flag_stride)) return;")
    out.append("    const eip::cptr<J3cData> jt = kJobs ? j3c_item_job(a, item) : (eip::cptr<J3cData>) nullptr;")
    out.append("#define JD(f) (kJobs ? (decltype(a.d.f)) jt->f : a.d.f)")
    out.append("    const int sp = eip::ro(a.items)[2 * item];")
    out.append("    const int ks_first = eip::ro(a.items)[2 * item + 1];")
    out.append("    const int ks_raw = ks_first + tid;")
    out.append("    const int ks_last = a.items_ks_end ? eip::ro(a.items_ks_end)[item] : (int) JD(n_ks);")
    out.append("    const bool active = ks_raw < ks_last;")
    out.append("    const int ks = active ? ks_raw : ks_last - 1;")
    out.append("    const int ksh = JD(ks_perm) ? eip::ro(JD(ks_perm))[ks] : ks;")
    out.append("    const int kp0 = eip::ro(JD(ks_prim_off))[ksh], kp1 = eip::ro(JD(ks_prim_off))[ksh + 1];")
    out.append("    const int pp0 = a.items_pp ? eip::ro(a.items_pp)[2 * item] : eip::ro(JD(sp_pp_off))[sp];")
    out.append("    const int pp1 = a.items_pp ? eip::ro(a.items_pp)[2 * item + 1] : eip::ro(JD(sp_pp_off))[sp + 1];")
    for ij in range(nij):
        for m in ms:
            out.append(f"    {ty} acc_{ij}_{m} = 0;")
    # Auxiliary primitives (per lane) outside, primitive pairs (wave-uniform) inside: the pair matrix
    # E of one primitive pair is then loaded right where it is consumed, as scalar loads, instead of
    # being hoisted out of an inner loop and spilled from SGPRs into VGPR lanes.
    stride = pair_stride(la, lb, spherical)
    esize = 4 if f32 else 8
    row_bytes = stride * esize
    # K consecutive rows per buffer (small classes): one wait per K primitive pairs instead of per pair
    K = max(1, min(STAGE_MAX_ROWS, STAGE_BATCH_BYTES // row_bytes))
    n_pieces = (K * row_bytes + 1023) // 1024              # 64 lanes x 16 bytes per global->LDS load
    if use_lds:
        ne = n_pieces * 1024 // esize
        out.append(f"    {ty} (* __restrict__ sE)[{ne}] = reinterpret_cast<{ty} (*)[{ne}]>(smem);")
        out.append(f"    double (* __restrict__ sP)[{4 * K}] = reinterpret_cast<double (*)[{4 * K}]>(smem + {2 * ne * esize});")
        out.append("    const int lane = tid;")
        out.append(f"    const char* __restrict__ Ebase = (const char*) JD({'E32' if f32 else 'E'});")
        out.append("    const char* __restrict__ Pbase = (const char*) JD(pp_P);")
        # stage(pp, nrows, buf): issue the asynchronous loads of rows [pp, pp + nrows) into buffer buf
        out.append("#define EIP_STAGE(PP, NR, BUF) do { \\")
        out.append(f"        const char* src_ = Ebase + (size_t) (PP) * {row_bytes} + 16 * lane; \\")
        out.append(f"        const int nbytes_ = (NR) * {row_bytes}; \\")
        for r in range(n_pieces):
            cond = "" if (r + 1) * 1024 <= row_bytes else f"if (16 * lane + {r * 1024} < nbytes_) "
            out.append(f"        {cond}__builtin_amdgcn_global_load_lds((const void*) (src_ + {r * 1024}), "
                       f"(__attribute__((address_space(3))) void*) ((char*) &sE[BUF][0] + {r * 1024}), 16, 0, 0); \\")
        out.append("        if (lane < 2 * (NR)) __builtin_amdgcn_global_load_lds((const void*) (Pbase + (size_t) (PP) * 32 + 16 * lane), "
                   "(__attribute__((address_space(3))) void*) &sP[BUF][0], 16, 0, 0); \\")
        out.append("    } while (0)")
    nv = (nnz + 15) // 16
    blds = boys_in_lds(la, lb, lc, precision, boys_lds)
    if blds:
        boff = j3c_smem_bytes(la, lb, lc, m0, m1, precision, stage, False)
        npc = (BOYS_LDS_BYTES + 1023) // 1024
        out.append(f"    const double* __restrict__ sB = reinterpret_cast<const double*>(smem + {boff});")
        out.append("    {   // the coarse Boys slice of order L into LDS (64 lanes x 16 bytes per copy), awaited at once")
        out.append(f"        const char* src_ = (const char*) (boys_lds64 + {L} * BOYS_LDS_STRIDE) + 16 * tid;")
        for r in range(npc):
            cond = "" if (r + 1) * 1024 <= BOYS_LDS_BYTES else f"if (16 * tid + {r * 1024} < {BOYS_LDS_BYTES}) "
            out.append(f"        {cond}__builtin_amdgcn_global_load_lds((const void*) (src_ + {r * 1024}), "
                       f"(__attribute__((address_space(3))) void*) (smem + {boff + r * 1024}), 16, 0, 0);")
        out.append("        __builtin_amdgcn_s_waitcnt(0x0F70);")
        out.append("        __syncthreads();")
        out.append("    }")
    if use_dpp:
        out.append("    const int l16 = tid & 15;")
        out.append(f"    const auto Erow = eip::ro(JD({'E32' if f32 else 'E'}));")
        out.append("    const auto Prow = eip::ro(JD(pp_P));")
        for v in range(nv):
            idx = f"{16 * v} + l16" if 16 * v + 15 < nnz else f"min({16 * v} + l16, {nnz - 1})"
            out.append(f"    const int ei_{v} = {idx};")
    if use_lds or use_dpp:
        # Every lane takes part in staging (DPP broadcasts need every lane active), so the loop over
        # auxiliary primitives runs to the wave's largest primitive count; lanes past their own count
        # contribute with coefficient 0.
        out.append("    const int nkp_wave = eip::wave_max(kp1 - kp0);")
        out.append("    for (int t_kp = 0; t_kp < nkp_wave; ++t_kp) {")
        out.append("      const bool kp_on = kp0 + t_kp < kp1;")
        out.append("      const int kp = kp_on ? kp0 + t_kp : kp1 - 1;")
        out.append("      const double Cx = eip::ro(JD(ka_C))[4 * kp + 0], Cy = eip::ro(JD(ka_C))[4 * kp + 1], Cz = eip::ro(JD(ka_C))[4 * kp + 2];")
        out.append("      const double q = eip::ro(JD(ka_C))[4 * kp + 3];")
        out.append("      const double ck = kp_on ? eip::ro(JD(ka_coef))[kp] : 0.0;")
        out.extend(_aux_prefactor_lines(lc))
    else:
        out.append("    for (int kp = kp0; kp < kp1; ++kp) {")
        out.append("      const double Cx = eip::ro(JD(ka_C))[4 * kp + 0], Cy = eip::ro(JD(ka_C))[4 * kp + 1], Cz = eip::ro(JD(ka_C))[4 * kp + 2];")
        out.append("      const double q = eip::ro(JD(ka_C))[4 * kp + 3];")
        out.append("      const double ck = eip::ro(JD(ka_coef))[kp];")
        out.extend(_aux_prefactor_lines(lc))
    if use_dpp:
        # Row and centre of the first primitive pair.  In the loop, each register Er_v is reloaded with
        # the next pair's entries right after the contraction's last use of it, so one set of
        # registers holds the row and the loads overlap the rest of the iteration.
        out.append(f"      {ty} " + ", ".join(f"Er_{v} = 0" for v in range(nv)) + ";")
        out.append("      double Pn0 = 0, Pn1 = 0, Pn2 = 0, Pn3 = 0;")
        out.append("      if (pp0 < pp1) {")
        for v in range(nv):
            out.append(f"        Er_{v} = Erow[(size_t) pp0 * {stride} + ei_{v}];")
        out.append("        Pn0 = Prow[4 * (size_t) pp0 + 0]; Pn1 = Prow[4 * (size_t) pp0 + 1]; Pn2 = Prow[4 * (size_t) pp0 + 2]; Pn3 = Prow[4 * (size_t) pp0 + 3];")
        out.append("      }")
        if prefetch:
            # prologue of the first pair and its Boys row; centre of the second
            out.append("      double Pm0 = 0, Pm1 = 0, Pm2 = 0, Pm3 = 0;")
            out.append("      double nzeta = 0, npf = 0, ndx = 0, ndy = 0, ndz = 0, nT = 0;")
            out.append(f"      double nrow[BOYS_K64 + 1];")
            out.append("      if (pp0 < pp1) {")
            out.append("        const int pq1 = min(pp0 + 1, pp1 - 1);")
            out.append("        Pm0 = Prow[4 * (size_t) pq1 + 0]; Pm1 = Prow[4 * (size_t) pq1 + 1]; Pm2 = Prow[4 * (size_t) pq1 + 2]; Pm3 = Prow[4 * (size_t) pq1 + 3];")
            out.extend(_prologue_lines("Pn", "        "))
            out.append(f"        eip::boys64_fetch<{L}>(nT, nrow);")
            out.append("      }")
        out.append("      for (int pp = pp0; pp < pp1; ++pp) {")
        if prefetch:
            out.append("        const double zeta = nzeta, pf = npf, dPQx = ndx, dPQy = ndy, dPQz = ndz, Tc = nT;")
            out.append("        double crow[BOYS_K64 + 1];")
            out.append("        for (int j = 0; j <= BOYS_K64; ++j) crow[j] = nrow[j];")
            out.append("        const int ppn = min(pp + 1, pp1 - 1);   // the last pair reloads its own row")
            out.append("        // the next pair (centre loaded one iteration ago): prologue and Boys row now, used next iteration")
            out.extend(_prologue_lines("Pm", "        "))
            out.append(f"        eip::boys64_fetch<{L}>(nT, nrow);")
            out.append("        { const int pq2 = min(pp + 2, pp1 - 1);")
            out.append("          Pm0 = Prow[4 * (size_t) pq2 + 0]; Pm1 = Prow[4 * (size_t) pq2 + 1]; Pm2 = Prow[4 * (size_t) pq2 + 2]; Pm3 = Prow[4 * (size_t) pq2 + 3]; }")
        else:
            out.append("        const double Px = Pn0, Py = Pn1, Pz = Pn2, p = Pn3;")
            out.append("        const int ppn = min(pp + 1, pp1 - 1);   // the last pair reloads its own row")
    elif use_lds and K == 1:
        out.append("      if (pp0 < pp1) EIP_STAGE(pp0, 1, 0);")
        out.append("      for (int pp = pp0, buf = 0; pp < pp1; ++pp, buf ^= 1) {")
        # vmcnt(0): the asynchronous global->LDS loads of row pp have landed (the compiler does not
        # track them across the loop back-edge); the barrier orders the LDS accesses.
        out.append("        __builtin_amdgcn_s_waitcnt(0x0F70);")
        out.append("        __syncthreads();")
        out.append(f"        const {ty}* __restrict__ E = &sE[buf][0];")
        out.append("        const double Px = sP[buf][0], Py = sP[buf][1], Pz = sP[buf][2], p = sP[buf][3];")
    elif use_lds:
        out.append(f"      if (pp0 < pp1) EIP_STAGE(pp0, min({K}, pp1 - pp0), 0);")
        out.append(f"      for (int b0 = pp0, buf = 0; b0 < pp1; b0 += {K}, buf ^= 1) {{")
        out.append("        __builtin_amdgcn_s_waitcnt(0x0F70);")
        out.append("        __syncthreads();")
        out.append(f"        const int nr = min({K}, pp1 - b0);")
        out.append("        for (int r = 0; r < nr; ++r) {")
        out.append(f"        const {ty}* __restrict__ E = &sE[buf][r * {stride}];")
        out.append("        const double Px = sP[buf][4 * r + 0], Py = sP[buf][4 * r + 1], Pz = sP[buf][4 * r + 2], p = sP[buf][4 * r + 3];")
    else:
        out.append("      for (int pp = pp0; pp < pp1; ++pp) {")
        out.append(f"        const auto E = eip::ro(JD({'E32' if f32 else 'E'})) + (size_t) pp * {stride};")
        out.append("        const auto Pp = eip::ro(JD(pp_P)) + 4 * (size_t) pp;\n        const double Px = Pp[0], Py = Pp[1], Pz = Pp[2], p = Pp[3];")
    if prefetch:
        out.append(f"        const double PQx = dPQx, PQy = dPQy, PQz = dPQz;")
        out.append(f"        const double T = Tc;")
        out.append(f"        double F[{L + 1}];")
        out.append(f"        eip::boys64_eval<{L}>(T, crow, F);")
    else:
        # geometry and prefactor in double always (cheap), then cast
        out.append("        const double dPQx = Px - Cx, dPQy = Py - Cy, dPQz = Pz - Cz;")
        # zeta = p q / (p + q) and pf = ck 2 pi^(5/2) hl / (p q sqrt(p + q)) from a reciprocal square
        # root and a reciprocal (hardware estimates refined by two Newton steps: a few ulp), not IEEE
        # divisions
        out.append("        double zeta, pf;")
        out.append("        if (a.omega == 0.0) {")
        out.append("            const double rs = eip::rsq_nr(p + q);")
        out.append("            zeta = p * q * (rs * rs);")
        out.append("            pf = ckq * eip::rcp_nr(p) * rs;")
        out.append("        } else {")
        # zeta = 1 / (1/p + 1/q + 1/w^2), pf = ck 2 pi^(5/2) hl w / (p q sqrt(p q + (p + q) w^2)) by the same
        # refined reciprocal and reciprocal square root (reference three_center.cpp:46-49)
        out.append("            const double w2 = a.omega * a.omega;")
        out.append("            zeta = p * q * eip::rcp_nr(q + p + p * q * eip::rcp_nr(w2));")
        out.append("            pf = ckq * a.omega * eip::rcp_nr(p) * eip::rsq_nr(p * q + (p + q) * w2);")
        out.append("        }")
        out.append(f"        const {ty} PQx = ({ty}) dPQx, PQy = ({ty}) dPQy, PQz = ({ty}) dPQz;")
        out.append(f"        const {ty} T = ({ty}) (zeta * (dPQx * dPQx + dPQy * dPQy + dPQz * dPQz));")
        out.append(f"        {ty} F[{L + 1}];")
        if blds:
            out.append(f"        eip::boys64_lds<{L}>(T, sB, F);")
        else:
            out.append(f"        eip::boys{'32' if f32 else '64'}<{L}>(T, F);")
    if prefetch:
        pass                      # centres are loaded two ahead above
    elif use_dpp:
        # centre of the next primitive pair (scalar loads), after the Boys table loads
        out.append("        Pn0 = Prow[4 * (size_t) ppn + 0]; Pn1 = Prow[4 * (size_t) ppn + 1]; Pn2 = Prow[4 * (size_t) ppn + 2]; Pn3 = Prow[4 * (size_t) ppn + 3];")
    elif use_lds and K == 1:
        out.append("        if (pp + 1 < pp1) EIP_STAGE(pp + 1, 1, buf ^ 1);   // after the Boys table loads: their wait excludes it")
    elif use_lds:
        out.append(f"        if (r == 0 && b0 + {K} < pp1) EIP_STAGE(b0 + {K}, min({K}, pp1 - b0 - {K}), buf ^ 1);")
    out.append(f"        const {ty} m2z = ({ty}) (-2.0 * zeta);")
    out.append(f"        const {ty} R000_0 = ({ty}) pf * F[0];")
    for n in range(1, L + 1):
        out.append(f"        const {ty} pf_{n} = {'(' + ty + ') pf' if n == 1 else f'pf_{n - 1}'} * m2z;")
        out.append(f"        const {ty} R000_{n} = pf_{n} * F[{n}];")
    rdemand = _RDemand(needed, ty) if r_order == "demand" else None
    if rdemand is None:
        _emit_R(out, L, needed, ty)
    # contraction, t-major
    k = 0
    if lds_dpp:
        # group g of 16 entries in register Eg_{g % 3}; groups 0 and 1 loaded here, group g + 2 when g starts
        out.append("        const int l16s = tid & 15;")
        out.append(f"        {ty} Eg_0 = E[min(l16s, {nnz - 1})];")
        out.append(f"        {ty} Eg_1 = E[min(16 + l16s, {nnz - 1})];" if nnz > 16 else f"        {ty} Eg_1 = 0;")
        out.append(f"        {ty} Eg_2 = 0;")
    for t in ts:
        vnames = {}
        for m in ms:
            r, terms = aux_rows[m]
            if rdemand is not None:
                for s in terms:
                    rdemand.request(out, (t[0] + s[0], t[1] + s[1], t[2] + s[2]))
            pieces = []
            for s, (qf, j) in terms.items():
                assert j == 0
                coef = SqrtRational(r, qf).to_float()
                rn = _r_name(0, (t[0] + s[0], t[1] + s[1], t[2] + s[2]))
                pieces.append(rn if coef == 1.0 else f"{lit(coef)} * {rn}")
            vn = f"V_{t[0]}_{t[1]}_{t[2]}_{m}"
            out.append(f"        const {ty} {vn} = {' + '.join(pieces)};")
            vnames[m] = vn
        while k < len(layout) and layout[k][0] == t:
            row = layout[k][1]
            if use_dpp:
                out.append(f"        {{ const {ty} e = eip::bcast16<{k % 16}>(Er_{k // 16});")
            elif lds_dpp:
                g = k // 16
                if k % 16 == 0 and 16 * (g + 2) < nnz:
                    out.append(f"        Eg_{(g + 2) % 3} = E[min({16 * (g + 2)} + l16s, {nnz - 1})];")
                out.append(f"        {{ const {ty} e = eip::bcast16<{k % 16}>(Eg_{g % 3});")
            else:
                out.append(f"        {{ const {ty} e = E[{k}];")
            for m in ms:
                out.append(f"          acc_{row}_{m} = fma{'f' if f32 else ''}(e, {vnames[m]}, acc_{row}_{m});")
            out.append("        }")
            if use_dpp and (k % 16 == 15 or k == nnz - 1):
                v = k // 16
                out.append(f"        Er_{v} = Erow[(size_t) ppn * {stride} + ei_{v}];")
            k += 1
    out.append("      }")
    if use_lds and K > 1:
        out.append("      }")
    out.append("    }")
    out.append("    if (!active) return;")
    out.append("    const int col = eip::ro(JD(ks_col))[ksh];")
    out.append("    const int I = eip::ro(JD(sp_i))[sp], J = eip::ro(JD(sp_j))[sp];")
    out.append("    __attribute__((address_space(1))) double* __restrict__ o = (__attribute__((address_space(1))) double*) JD(out) + (JD(sp_out_base) ? eip::ro(JD(sp_out_base))[sp] : 0LL);")
    out.append("    const long long ld = JD(sp_ld) ? (long long) eip::ro(JD(sp_ld))[sp] : JD(ld);")
    out.append("    if (a.accumulate == 0) {")
    out.append("     if (a.layout == 0) {")
    for ij in range(nij):
        mi, mj = divmod(ij, nj)
        out.append(f"      if (I != J || {mi} >= {mj}) {{ const long long row = eip::packed_row(I + {mi}, J + {mj}) * ld + col;")
        for m in ms:
            out.append(f"        o[row + {m}] = (double) acc_{ij}_{m};")
        out.append("      }")
    out.append("     } else {")
    out.append(f"      const long long base = (long long) sp * {nij};")
    for ij in range(nij):
        out.append(f"      {{ const long long row = (base + {ij}) * ld + col;")
        for m in ms:
            out.append(f"        o[row + {m}] = (double) acc_{ij}_{m};")
        out.append("      }")
    out.append("     }")
    out.append("    } else {")
    # accumulate: loads of a group of outputs are issued together, then the sums are stored
    out.append(f"      const long long base = (long long) sp * {nij};")
    pairs_ = [(ij, m) for ij in range(nij) for m in ms]
    for ij in range(nij):
        mi, mj = divmod(ij, nj)
        out.append(f"      const bool ok_{ij} = a.layout != 0 || I != J || {mi} >= {mj};")
        out.append(f"      const long long row_{ij} = (a.layout == 0 ? eip::packed_row(I + {mi}, J + {mj}) : base + {ij}) * ld + col;")
    for g0 in range(0, len(pairs_), 16):
        grp = pairs_[g0:g0 + 16]
        for ij, m in grp:
            out.append(f"      const double o_{ij}_{m} = ok_{ij} ? o[row_{ij} + {m}] : 0.0;")
        for ij, m in grp:
            out.append(f"      if (ok_{ij}) o[row_{ij} + {m}] = o_{ij}_{m} + (double) acc_{ij}_{m};")
    out.append("    }")
    out.append("#undef JD")
    if use_lds:
        out.append("#undef EIP_STAGE")
    out.append("}")
    smem_bytes = j3c_smem_bytes(la, lb, lc, m0, m1, precision, stage, boys_lds)
    # the persistent queue kernels include this source for its bodies only
    out.append("#ifndef EIP_BODIES_ONLY")
    for kname, jobs in ((name, "false"), (name + "_jobs", "true")):
        if smem_bytes:
            out.append(f'extern "C" __global__ void __launch_bounds__(64) {kname}(const J3cArgs a) {{ '
                       f'__shared__ __attribute__((aligned(16))) char smem[{smem_bytes}]; '
                       f'{name}_body<{jobs}>(a, eip::xcd_remap(blockIdx.x, gridDim.x), smem, (int) threadIdx.x); }}')
        else:
            out.append(f'extern "C" __global__ void __launch_bounds__(64) {kname}(const J3cArgs a) {{ '
                       f'{name}_body<{jobs}>(a, eip::xcd_remap(blockIdx.x, gridDim.x), nullptr, (int) threadIdx.x); }}')
    out.append("#endif")
    return name, "\n".join(out)


def j3c_smem_bytes(la, lb, lc, m0, m1, precision="f64", stage=None, boys_lds=False):
    """LDS bytes the vector three-centre body of this class chunk stages its pair-matrix rows in (two
    buffers of K rows and their centres; 0 for DPP-staged bodies), followed by the coarse Boys slice when
    boys_lds applies to the class."""
    stage = stage or STAGE
    extra = BOYS_LDS_BYTES if boys_in_lds(la, lb, lc, precision, boys_lds) else 0
    use_dpp = stage == "dpp" and pair_nnz(la, lb, True) <= DPP_MAX_ENTRIES
    if use_dpp or not STAGE_LDS or stage == "sgpr":
        return extra
    esize = 4 if precision == "f32" else 8
    row_bytes = pair_stride(la, lb, True) * esize
    K = max(1, min(STAGE_MAX_ROWS, STAGE_BATCH_BYTES // row_bytes))
    n_pieces = (K * row_bytes + 1023) // 1024
    return 2 * n_pieces * 1024 + 2 * 4 * K * 8 + extra


# ---------------------------------------------------------------------------------------------
# Nuclear attraction kernel: lanes over point charges
# ---------------------------------------------------------------------------------------------

def emit_nuc(la, lb):
    spherical = True
    name = f"v_{la}{lb}"
    L = la + lb
    layout = pair_layout(la, lb, spherical)
    nnz = len(layout)
    nij = n_out(la, spherical) * n_out(lb, spherical)
    ts = []
    for t, _ in layout:
        if t not in ts:
            ts.append(t)
    out = [f"// nuclear attraction ({la}{lb}): {nnz} pair entries"]
    out.append(f'extern "C" __global__ void __launch_bounds__(64) {name}(const NucArgs a)')
    out.append("{")
    out.append("    const int blk = eip::xcd_remap(blockIdx.x, gridDim.x);")
    out.append("    const int nchunk = (a.n_charge + 63) / 64;")
    out.append("    const int sp = blk / nchunk;")
    out.append("    if (sp >= a.n_sp) return;")
    out.append("    const int c_raw = (blk - sp * nchunk) * 64 + (int) threadIdx.x;")
    out.append("    const bool active = c_raw < a.n_charge;")
    out.append("    const int ic = active ? c_raw : a.n_charge - 1;")
    out.append("    const double Cx = a.charge[3 * ic + 0], Cy = a.charge[3 * ic + 1], Cz = a.charge[3 * ic + 2];")
    out.append("    const int pp0 = a.sp_pp_off[sp], pp1 = a.sp_pp_off[sp + 1];")
    for ij in range(nij):
        out.append(f"    double acc_{ij} = 0;")
    out.append("    for (int pp = pp0; pp < pp1; ++pp) {")
    out.append(f"        const double* __restrict__ E = a.E + (size_t) pp * {pair_stride(la, lb, spherical)};")
    out.append("        const double p = a.pp_P[4 * pp + 3];")
    out.append("        const double PQx = a.pp_P[4 * pp + 0] - Cx, PQy = a.pp_P[4 * pp + 1] - Cy, PQz = a.pp_P[4 * pp + 2] - Cz;")
    out.append("        double zeta, pf;")
    out.append("        if (a.omega == 0.0) { zeta = p; pf = eip::kTwoPi / p; }")
    out.append("        else { const double w2 = a.omega * a.omega; zeta = p * w2 / (p + w2); pf = eip::kTwoPi * a.omega / (p * sqrt(p + w2)); }")
    out.append("        const double T = zeta * (PQx * PQx + PQy * PQy + PQz * PQz);")
    out.append(f"        double F[{L + 1}];")
    out.append(f"        eip::boys64<{L}>(T, F);")
    out.append("        const double m2z = -2.0 * zeta;")
    out.append("        const double R000_0 = pf * F[0];")
    for n in range(1, L + 1):
        out.append(f"        const double pf_{n} = {'pf' if n == 1 else f'pf_{n - 1}'} * m2z;")
        out.append(f"        const double R000_{n} = pf_{n} * F[{n}];")
    _emit_R(out, L, set(ts), "double")
    for k, (t, row) in enumerate(layout):
        out.append(f"        acc_{row} = fma(E[{k}], {_r_name(0, t)}, acc_{row});")
    out.append("    }")
    out.append("    if (!active) return;")
    out.append("    const long long base = (long long) sp * %d;" % nij)
    for ij in range(nij):
        out.append(f"    a.out[(base + {ij}) * a.ld + ic] = acc_{ij};")
    out.append("}")
    return name, "\n".join(out)


# ---------------------------------------------------------------------------------------------
# Two-centre kernel: lanes over auxiliary shells J for one shell I per wave
# ---------------------------------------------------------------------------------------------

def emit_j2c(li, lj):
    spherical = True
    name = f"j2c_{li}{lj}"
    L = li + lj
    ri = aux_transform(li, spherical)
    rj = aux_transform(lj, spherical)
    ni, nj = len(ri), len(rj)
    out = [f"// two-centre ({li}|{lj})"]
    out.append(f'extern "C" __global__ void __launch_bounds__(64) {name}(const J2cArgs a)')
    out.append("{")
    out.append("    const int blk = eip::xcd_remap(blockIdx.x, gridDim.x);")
    out.append("    const int nchunk = (a.n_j + 63) / 64;")
    out.append("    const int si = blk / nchunk;")
    out.append("    if (si >= a.n_i) return;")
    out.append("    const int sj_raw = (blk - si * nchunk) * 64 + (int) threadIdx.x;")
    out.append("    const bool active = sj_raw < a.n_j;")
    out.append("    const int sj = active ? sj_raw : a.n_j - 1;")
    out.append("    const int ip0 = a.i_prim_off[si], ip1 = a.i_prim_off[si + 1];")
    out.append("    const int jp0 = a.j_prim_off[sj], jp1 = a.j_prim_off[sj + 1];")
    for mi in range(ni):
        for mj in range(nj):
            out.append(f"    double acc_{mi}_{mj} = 0;")
    out.append("    for (int ip = ip0; ip < ip1; ++ip) {")
    out.append("      const double p = a.i_A[4 * ip + 3], ci = a.i_coef[ip];")
    out.append("      const double Ax = a.i_A[4 * ip + 0], Ay = a.i_A[4 * ip + 1], Az = a.i_A[4 * ip + 2];")
    out.append("      for (int jp = jp0; jp < jp1; ++jp) {")
    out.append("        const double q = a.j_A[4 * jp + 3];")
    out.append("        const double PQx = Ax - a.j_A[4 * jp + 0], PQy = Ay - a.j_A[4 * jp + 1], PQz = Az - a.j_A[4 * jp + 2];")
    out.append("        const double coef = ci * a.j_coef[jp];")
    out.append("        double zeta, pf;")
    out.append("        if (a.omega == 0.0) { zeta = p * q / (p + q); pf = coef * eip::kTwoPi25 / (p * q * sqrt(p + q)); }")
    out.append("        else { const double w2 = a.omega * a.omega; zeta = 1.0 / (1.0 / p + 1.0 / q + 1.0 / w2);")
    out.append("               pf = coef * eip::kTwoPi25 * a.omega / (p * q * sqrt(p * q + (p + q) * w2)); }")
    for side, l, ex in (("i", li, "p"), ("j", lj, "q")):
        if l > 0:
            out.append(f"        {{ const double h = 0.5 / {ex}; double hp = h;")
            for _ in range(l - 1):
                out.append("          hp *= h;")
            sign = "-" if (side == "j" and l % 2) else ""
            out.append(f"          pf *= {sign}hp; }}")
    out.append("        const double T = zeta * (PQx * PQx + PQy * PQy + PQz * PQz);")
    out.append(f"        double F[{L + 1}];")
    out.append(f"        eip::boys64<{L}>(T, F);")
    out.append("        const double m2z = -2.0 * zeta;")
    out.append("        const double R000_0 = pf * F[0];")
    for n in range(1, L + 1):
        out.append(f"        const double pf_{n} = {'pf' if n == 1 else f'pf_{n - 1}'} * m2z;")
        out.append(f"        const double R000_{n} = pf_{n} * F[{n}];")
    needed = set()
    for _, ti in ri:
        for t in ti:
            for _, sj in rj:
                for s in sj:
                    needed.add((t[0] + s[0], t[1] + s[1], t[2] + s[2]))
    _emit_R(out, L, needed, "double")
    # Two stages: W_mj(t) = sum_s c_j(mj, s) R_{t+s} for every Hermite t of the i side, then
    # acc_mi,mj += sum_t c_i(mi, t) W_mj(t).  Only the two sides' own transform coefficients appear (a
    # few dozen constants), not their products -- those would all be hoisted out of the primitive loops
    # into SGPRs -- and the flop count drops from sum_mi,mj |t_i| |s_j| to (|T_i| + n_i) per mj.
    ts_i = sorted({t for _, ti in ri for t in ti}, key=lambda t: (sum(t), -t[0], -t[1]))
    for mj, (r2, t2) in enumerate(rj):
        for t in ts_i:
            pieces = []
            for s, (q2, j2) in t2.items():
                assert j2 == 0
                coef = SqrtRational(r2, q2).to_float()
                rn = _r_name(0, (t[0] + s[0], t[1] + s[1], t[2] + s[2]))
                pieces.append(rn if coef == 1.0 else f"{fmt(coef)} * {rn}")
            out.append(f"        const double W_{mj}_{t[0]}_{t[1]}_{t[2]} = {' + '.join(pieces)};")
        for mi, (r1, t1) in enumerate(ri):
            pieces = []
            for t, (q1, j1) in t1.items():
                assert j1 == 0
                coef = SqrtRational(r1, q1).to_float()
                wn = f"W_{mj}_{t[0]}_{t[1]}_{t[2]}"
                pieces.append(wn if coef == 1.0 else f"{fmt(coef)} * {wn}")
            out.append(f"        acc_{mi}_{mj} += {' + '.join(pieces)};")
    out.append("      }")
    out.append("    }")
    out.append("    if (!active) return;")
    out.append("    const int I = a.i_col[si], J = a.j_col[sj];")
    out.append("    if (a.same_class && I > J) return;   // reference two_center.cpp:138")
    out.append("    double* __restrict__ o = a.out;")
    for mi in range(ni):
        for mj in range(nj):
            out.append(f"    o[(long long) (I + {mi}) * a.ld + J + {mj}] = acc_{mi}_{mj};")
    out.append("    if (I != J && a.mirror) {")
    for mi in range(ni):
        for mj in range(nj):
            out.append(f"      o[(long long) (J + {mj}) * a.ld + I + {mi}] = acc_{mi}_{mj};")
    out.append("    }")
    out.append("}")
    return name, "\n".join(out)


# ---------------------------------------------------------------------------------------------
# Work split of a class over kernels (register budget)
# ---------------------------------------------------------------------------------------------

def _needed_R(la, lb, lc, m0, m1):
    """Number of Hermite integrals a chunk [m0, m1) of class (la lb|lc) keeps live."""
    layout = pair_layout(la, lb, True)
    ts = {t for t, _ in layout}
    rows = aux_transform(lc, True)
    need = set()
    for m in range(m0, m1):
        for s in rows[m][1]:
            for t in ts:
                need.add((t[0] + s[0], t[1] + s[1], t[2] + s[2]))
    return len(need)


# Register budget per lane, in doubles: accumulators + live Hermite integrals + working set must stay
# well inside the 512 registers (256 VGPR + 256 AGPR) of a wave at occupancy 1, so no kernel spills to
# scratch memory.  'shrink' lets the build lower the budget of classes whose kernels still spilled.
REG_BUDGET_DOUBLES = 200


def m_chunks(la, lb, lc, shrink=0, budget=None):
    """Split the 2lc+1 auxiliary components of class (la lb|lc) over kernels to bound registers."""
    nij = (2 * la + 1) * (2 * lb + 1)
    nk = 2 * lc + 1
    budget = (budget if budget is not None else REG_BUDGET_DOUBLES) >> shrink
    per = nk
    while per > 1:
        n = -(-nk // per)
        p = -(-nk // n)
        worst = max(nij * (min(nk, m + p) - m) + _needed_R(la, lb, lc, m, min(nk, m + p))
                    for m in range(0, nk, p))
        if worst <= budget:
            break
        per -= 1
    n = -(-nk // per)
    per = -(-nk // n)          # balance
    return [(m, min(nk, m + per)) for m in range(0, nk, per)]


# ---------------------------------------------------------------------------------------------
# Three-centre kernel on the FP64 matrix cores (v_mfma_f64_16x16x4f64)
# ---------------------------------------------------------------------------------------------
#
# For one auxiliary component m and one block b of 16 pair-output rows, the contraction over a group
# of four primitive pairs is a 16x16x4 matrix product
#     acc_m,b[row, aux] += sum_{k < 4} E^{pp_k}[16 b + row, t] * V^{pp_k, aux}[t, m]
# issued once per pair Hermite index t.  Lane l of a wave holds A[row = l % 16][k = l / 16] (its own
# primitive pair's pair-matrix entry, read from the dense layout below), B[k = l / 16][aux = l % 16]
# (V of its own triplet: primitive pair pp_{l/16} with auxiliary shell aux_{l%16}, computed in the
# lane's registers like the vector kernels do) and, in accumulator register r < 4, row 4 r + l / 16
# of column l % 16 (lane layout checked on the device by a probe of the instruction).  A workgroup of four waves takes one work item (shell pair, 64 auxiliary shells) of
# the vector kernels' item tables, 16 auxiliary shells per wave.  The accumulators live in the matrix
# cores' registers, so a chunk of auxiliary components is bounded by the live Hermite integrals only.

MFMA_WAVES = 4                   # waves per workgroup (16 auxiliary shells each)
# Pair-matrix operands of a group loaded at the top of its iteration, ahead of the Boys function and
# the Hermite integrals, so their latency is hidden behind that work (the rest are loaded in the
# contraction, a few ahead of use): as many as fit beside the live Hermite integrals in this many
# vector registers, at most MFMA_E_PREFETCH_MAX.
MFMA_E_PREFETCH_VGPRS = 216
MFMA_E_PREFETCH_MAX = 24


def mfma_blocks(la, lb):
    """Blocks of 16 pair-output rows."""
    return -(-(n_out(la, True) * n_out(lb, True)) // 16)


def mfma_hermite(la, lb):
    """Pair Hermite indices t in the order of the packed layout (and of the dense layout's t axis)."""
    ts = []
    for t, _ in pair_layout(la, lb, True):
        if t not in ts:
            ts.append(t)
    return ts


def mfma_dense_width(la, lb):
    """Doubles per primitive pair in the dense layout: [t][16 * blocks] (structural zeros stored)."""
    return len(mfma_hermite(la, lb)) * 16 * mfma_blocks(la, lb)


def emit_pair_dense(la, lb):
    """pairD_<la><lb>: packed pair-matrix rows -> the dense [pp][t][16 * blocks] layout the matrix-core
    kernels read (one thread per primitive pair)."""
    name = f"pairD_{la}{lb}"
    ts = mfma_hermite(la, lb)
    tix = {t: i for i, t in enumerate(ts)}
    W = 16 * mfma_blocks(la, lb)
    layout = pair_layout(la, lb, True)
    stride = pair_stride(la, lb, True)
    where = {(tix[t], row): k for k, (t, row) in enumerate(layout)}
    L = [f"// dense pair matrix of class ({la},{lb}): {len(ts)} Hermite indices x {W} rows"]
    L.append(f'extern "C" __global__ void __launch_bounds__(256) {name}(const double* __restrict__ E, '
             f'double* __restrict__ D, int n_pp)')
    L.append("{")
    L.append("    const int pp = blockIdx.x * blockDim.x + threadIdx.x;")
    L.append("    if (pp >= n_pp) return;")
    L.append(f"    const double* __restrict__ e = E + (size_t) pp * {stride};")
    L.append(f"    double2* __restrict__ d = reinterpret_cast<double2*>(D + (size_t) pp * {len(ts) * W});")
    for ti in range(len(ts)):
        for r in range(0, W, 2):
            v = [f"e[{where[(ti, rr)]}]" if (ti, rr) in where else "0.0" for rr in (r, r + 1)]
            L.append(f"    d[{(ti * W + r) // 2}] = make_double2({v[0]}, {v[1]});")
    L.append("}")
    return name, "\n".join(L)


def mfma_chunks(la, lb, lc, r_budget, acc_budget=224):
    """Split the auxiliary components over matrix-core kernels: each chunk keeps at most r_budget
    Hermite integrals live (vector registers) and 8 * components * blocks accumulator registers."""
    nk = 2 * lc + 1
    nb = mfma_blocks(la, lb)
    per = nk
    while per > 1:
        n = -(-nk // per)
        p = -(-nk // n)
        ok = all(_needed_R(la, lb, lc, m, min(nk, m + p)) <= r_budget and 8 * (min(nk, m + p) - m) * nb <= acc_budget
                 for m in range(0, nk, p))
        if ok:
            break
        per -= 1
    n = -(-nk // per)
    per = -(-nk // n)
    return [(m, min(nk, m + per)) for m in range(0, nk, per)]


def emit_j3c_mfma(la, lb, lc, m0, m1):
    """Float64 three-centre kernel on the matrix cores for auxiliary components [m0, m1) of class
    (la lb | lc); reads the dense pair matrix (J3cData.E points to it) and the vector kernels' items."""
    spherical = True
    name = f"j3cm_{la}{lb}{lc}_{m0}_{m1}"
    L = la + lb + lc
    nij = n_out(la, spherical) * n_out(lb, spherical)
    nj = n_out(lb, spherical)
    nb = mfma_blocks(la, lb)
    W = 16 * nb
    ts = mfma_hermite(la, lb)
    DS = len(ts) * W
    aux_rows = aux_transform(lc, spherical)
    ms = list(range(m0, m1))
    needed = set()
    for t in ts:
        for m in ms:
            for s in aux_rows[m][1]:
                needed.add((t[0] + s[0], t[1] + s[1], t[2] + s[2]))
    out = [f"// ({la}{lb}|{lc}) aux components [{m0},{m1}) f64 matrix cores: L = {L}, {len(ts)} Hermite indices, "
           f"{nb} row blocks"]
    out.append("template <bool kJobs>")
    out.append(f"__device__ __forceinline__ void {name}_body(const J3cArgs a)")
    out.append("{")
    out.append("    const int item = a.item_begin + eip::xcd_remap(blockIdx.x, gridDim.x);")
    out.append("    if (item >= a.item_end) return;")
    out.append("    if (a.flags_in && !eip::item_flagged(a.flags_in + (size_t) item * a.flag_stride, a.flag_stride)) return;")
    out.append("    const eip::cptr<J3cData> jt = kJobs ? j3c_item_job(a, item) : (eip::cptr<J3cData>) nullptr;")
    out.append("#define JD(f) (kJobs ? (decltype(a.d.f)) jt->f : a.d.f)")
    out.append("    const int wave = (int) threadIdx.x >> 6, lane = (int) threadIdx.x & 63;")
    out.append("    const int kslot = lane >> 4, jcol = lane & 15;")
    out.append("    const int sp = eip::ro(a.items)[2 * item];")
    out.append("    const int ks_first = eip::ro(a.items)[2 * item + 1] + 16 * wave;")
    out.append("    const int ks_last = a.items_ks_end ? eip::ro(a.items_ks_end)[item] : (int) JD(n_ks);")
    out.append("    if (ks_first >= ks_last) return;                 // no auxiliary shell for this wave")
    out.append("    const int ks_raw = ks_first + jcol;")
    out.append("    const bool active = ks_raw < ks_last;")
    out.append("    const int ks = active ? ks_raw : ks_last - 1;")
    out.append("    const int ksh = JD(ks_perm) ? eip::ro(JD(ks_perm))[ks] : ks;")
    out.append("    const int kp0 = eip::ro(JD(ks_prim_off))[ksh], kp1 = eip::ro(JD(ks_prim_off))[ksh + 1];")
    out.append("    const int pp0 = a.items_pp ? eip::ro(a.items_pp)[2 * item] : eip::ro(JD(sp_pp_off))[sp];")
    out.append("    const int pp1 = a.items_pp ? eip::ro(a.items_pp)[2 * item + 1] : eip::ro(JD(sp_pp_off))[sp + 1];")
    out.append("    const double* __restrict__ ED = JD(E);")
    out.append("    const double* __restrict__ PP = JD(pp_P);")
    for m in ms:
        for b in range(nb):
            out.append(f"    eip::v4d acc_{m}_{b} = {{0.0, 0.0, 0.0, 0.0}};")
    out.append("    const int nkp_wave = eip::wave_max(kp1 - kp0);")
    out.append("    for (int t_kp = 0; t_kp < nkp_wave; ++t_kp) {")
    out.append("      const bool kp_on = kp0 + t_kp < kp1;")
    out.append("      const int kp = kp_on ? kp0 + t_kp : kp1 - 1;")
    out.append("      const double Cx = JD(ka_C)[4 * kp + 0], Cy = JD(ka_C)[4 * kp + 1], Cz = JD(ka_C)[4 * kp + 2];")
    out.append("      const double q = JD(ka_C)[4 * kp + 3];")
    out.append("      const double ck = kp_on ? JD(ka_coef)[kp] : 0.0;")
    out.extend(_aux_prefactor_lines(lc))
    # centre and exponent of the lane's primitive pair in the first group, then one group ahead
    out.append("      double Pn0 = 0, Pn1 = 0, Pn2 = 0, Pn3 = 1;")
    out.append("      if (pp0 < pp1) { const int q0 = min(pp0 + kslot, pp1 - 1);")
    out.append("        Pn0 = PP[4 * (size_t) q0 + 0]; Pn1 = PP[4 * (size_t) q0 + 1]; Pn2 = PP[4 * (size_t) q0 + 2]; Pn3 = PP[4 * (size_t) q0 + 3]; }")
    out.append("      for (int g0 = pp0; g0 < pp1; g0 += 4) {")
    out.append("        const int pq = g0 + kslot;")
    out.append("        const bool pp_on = pq < pp1;")
    out.append("        const int pp = pp_on ? pq : pp1 - 1;")
    out.append("        const double Px = Pn0, Py = Pn1, Pz = Pn2, p = Pn3;")
    out.append(f"        const double* __restrict__ Ed = ED + (size_t) pp * {DS} + jcol;")
    tb = [(ti, b) for ti in range(len(ts)) for b in range(nb)]
    n_pre = max(0, min(MFMA_E_PREFETCH_MAX, (MFMA_E_PREFETCH_VGPRS - 2 * len(needed)) // 2))
    pre = tb[:n_pre]
    for ti, b in pre:
        out.append(f"        const double ea_{ti}_{b} = Ed[{ti * W + 16 * b}];")
    out.append("        { const int qn = min(g0 + 4 + kslot, pp1 - 1);      // the next group's pair (the last reloads one)")
    out.append("          Pn0 = PP[4 * (size_t) qn + 0]; Pn1 = PP[4 * (size_t) qn + 1]; Pn2 = PP[4 * (size_t) qn + 2]; Pn3 = PP[4 * (size_t) qn + 3]; }")
    out.append("        const double dPQx = Px - Cx, dPQy = Py - Cy, dPQz = Pz - Cz;")
    out.append("        double zeta, pf;")
    out.append("        if (a.omega == 0.0) {")
    out.append("            const double rs = eip::rsq_nr(p + q);")
    out.append("            zeta = p * q * (rs * rs);")
    out.append("            pf = ckq * eip::rcp_nr(p) * rs;")
    out.append("        } else {")
    out.append("            const double w2 = a.omega * a.omega;")
    out.append("            zeta = p * q * eip::rcp_nr(q + p + p * q * eip::rcp_nr(w2));")
    out.append("            pf = ckq * a.omega * eip::rcp_nr(p) * eip::rsq_nr(p * q + (p + q) * w2);")
    out.append("        }")
    out.append("        if (!pp_on) pf = 0.0;                        // an empty slot of the group adds nothing")
    out.append("        const double PQx = dPQx, PQy = dPQy, PQz = dPQz;")
    out.append("        const double T = zeta * (dPQx * dPQx + dPQy * dPQy + dPQz * dPQz);")
    out.append(f"        double F[{L + 1}];")
    out.append(f"        eip::boys64<{L}>(T, F);")
    out.append("        const double m2z = -2.0 * zeta;")
    out.append("        const double R000_0 = pf * F[0];")
    for n in range(1, L + 1):
        out.append(f"        const double pf_{n} = {'pf' if n == 1 else f'pf_{n - 1}'} * m2z;")
        out.append(f"        const double R000_{n} = pf_{n} * F[{n}];")
    # level order here: with the Hermite integrals emitted next to their use, the VALU work interleaves
    # with the matrix-core instructions and the (ff|p), (ff|d) kernels ran 1.7-3.6x slower
    _emit_R(out, L, needed, "double")
    for ti, t in enumerate(ts):
        for m in ms:
            r, terms = aux_rows[m]
            pieces = []
            for s, (qf, j) in terms.items():
                assert j == 0
                coef = SqrtRational(r, qf).to_float()
                rn = _r_name(0, (t[0] + s[0], t[1] + s[1], t[2] + s[2]))
                pieces.append(rn if coef == 1.0 else f"{fmt(coef)} * {rn}")
            out.append(f"        const double V_{ti}_{m} = {' + '.join(pieces)};")
        for b in range(nb):
            if (ti, b) in pre:
                out.append(f"        {{ const double ea = ea_{ti}_{b};")
            else:
                out.append(f"        {{ const double ea = Ed[{ti * W + 16 * b}];")
            for m in ms:
                out.append(f"          acc_{m}_{b} = __builtin_amdgcn_mfma_f64_16x16x4f64(ea, V_{ti}_{m}, acc_{m}_{b}, 0, 0, 0);")
            out.append("        }")
    out.append("      }")
    out.append("    }")
    out.append("    if (!active) return;")
    out.append("    const int col = eip::ro(JD(ks_col))[ksh];")
    out.append("    const int I = eip::ro(JD(sp_i))[sp], J = eip::ro(JD(sp_j))[sp];")
    out.append("    __attribute__((address_space(1))) double* __restrict__ o = (__attribute__((address_space(1))) double*) JD(out) + (JD(sp_out_base) ? eip::ro(JD(sp_out_base))[sp] : 0LL);")
    out.append("    const long long ld = JD(sp_ld) ? (long long) eip::ro(JD(sp_ld))[sp] : JD(ld);")
    out.append(f"    const long long base = (long long) sp * {nij};")
    for b in range(nb):
        for r_ in range(4):
            out.append(f"    {{ const int ij = {16 * b + 4 * r_} + kslot;")
            out.append(f"      const int mi = ij / {nj}, mj = ij - mi * {nj};")
            out.append(f"      if (ij < {nij} && (a.layout != 0 || I != J || mi >= mj)) {{")
            out.append("        const long long row = (a.layout == 0 ? eip::packed_row(I + mi, J + mj) : base + ij) * ld + col;")
            for m in ms:
                out.append(f"        if (a.accumulate) o[row + {m}] += acc_{m}_{b}[{r_}]; else o[row + {m}] = acc_{m}_{b}[{r_}];")
            out.append("      }")
            out.append("    }")
    out.append("#undef JD")
    out.append("}")
    out.append(f'extern "C" __global__ void __launch_bounds__({64 * MFMA_WAVES}) {name}(const J3cArgs a) {{ {name}_body<false>(a); }}')
    out.append(f'extern "C" __global__ void __launch_bounds__({64 * MFMA_WAVES}) {name}_jobs(const J3cArgs a) {{ {name}_body<true>(a); }}')
    return name, "\n".join(out)


# ---------------------------------------------------------------------------------------------
# Float32-tier three-centre kernel on the bf16 matrix cores (v_mfma_f32_16x16x32_bf16)
# ---------------------------------------------------------------------------------------------
#
# The same work decomposition as the float64 matrix-core kernel (four primitive-pair slots x 16
# auxiliary shells per wave, four waves per item), with K = 32 = 4 pair slots x 8 consecutive Hermite
# indices per instruction: lane l holds A[row = l % 16][k = 8 (l / 16) + v] and B[k][aux = l % 16] for
# v < 8, and accumulator register r holds row 4 (l / 16) + r (checked on the device by a probe).  Every
# operand is split into BF16_PIECES bfloat16 pieces x = x_0 + x_1 + ... (x_0 = bf16(x), x_1 =
# bf16(x - x_0), ...); the products of pieces i, j with i + j < BF16_PIECES are accumulated in float32.
# With three pieces each operand is represented to 2^-24 of its magnitude and the dropped products are
# of that order, so a product carries float32-level relative error; the Boys function, the Hermite
# integrals and the auxiliary transform are evaluated in float32 as in the vector float32 kernels.

BF16_PIECES = 3
# steps (Hermite group, row block) the pair-matrix operand loads run ahead of their use
BF16_E_AHEAD = 2


def bf16_groups(la, lb):
    """Groups of 8 pair Hermite indices (the K slices of one lane)."""
    return -(-len(mfma_hermite(la, lb)) // 8)


def bf16_dense_width(la, lb):
    """bfloat16 values per primitive pair of the split dense layout [piece][group][block][row][8]."""
    return BF16_PIECES * bf16_groups(la, lb) * mfma_blocks(la, lb) * 16 * 8


def bf16_chunks(la, lb, lc, budget):
    """Split the auxiliary components over bf16 matrix-core kernels so that each chunk's live vector
    registers -- float Hermite integrals, the split auxiliary operands (4 registers per piece and
    component) and the accumulators (4 per component and row block) -- stay within budget."""
    nk = 2 * lc + 1
    nb = mfma_blocks(la, lb)

    def regs(m0, m1):
        n = m1 - m0
        return _needed_R(la, lb, lc, m0, m1) + 4 * BF16_PIECES * n + 4 * n * nb + 4 * BF16_PIECES + 48

    for n in range(1, nk + 1):
        per = -(-nk // n)
        plan = [(m, min(nk, m + per)) for m in range(0, nk, per)]
        if all(regs(m0, m1) <= budget for m0, m1 in plan):
            return plan
    return [(m, m + 1) for m in range(nk)]


def emit_pair_bf16(la, lb):
    """pairB_<la><lb>: packed float64 pair-matrix rows -> the split bfloat16 dense layout (one thread per
    primitive pair; 16-byte stores of 8 Hermite indices of one row)."""
    name = f"pairB_{la}{lb}"
    ts = mfma_hermite(la, lb)
    tix = {t: i for i, t in enumerate(ts)}
    nq, nb = bf16_groups(la, lb), mfma_blocks(la, lb)
    layout = pair_layout(la, lb, True)
    stride = pair_stride(la, lb, True)
    where = {(tix[t], row): k for k, (t, row) in enumerate(layout)}
    W = bf16_dense_width(la, lb)
    L = [f"// split bfloat16 dense pair matrix of class ({la},{lb}): {BF16_PIECES} pieces x {nq} groups x {nb} blocks"]
    L.append(f'extern "C" __global__ void __launch_bounds__(256) {name}(const double* __restrict__ E, '
             f'eip::bf8* __restrict__ D, int n_pp)')
    L.append("{")
    L.append("    const int pp = blockIdx.x * blockDim.x + threadIdx.x;")
    L.append("    if (pp >= n_pp) return;")
    L.append(f"    const double* __restrict__ e = E + (size_t) pp * {stride};")
    L.append(f"    eip::bf8* __restrict__ d = D + (size_t) pp * {W // 8};")
    for q in range(nq):
        for b in range(nb):
            for row in range(16):
                vals = []
                for v in range(8):
                    key = (8 * q + v, 16 * b + row)
                    vals.append(f"e[{where[key]}]" if key in where else "0.0")
                L.append("    { const double x[8] = {" + ", ".join(vals) + "};")
                L.append(f"      eip::bf8 h[{BF16_PIECES}];")
                L.append(f"      eip::split_bf16<{BF16_PIECES}>(x, h);")
                for pc in range(BF16_PIECES):
                    off = ((pc * nq + q) * nb + b) * 16 + row
                    L.append(f"      d[{off}] = h[{pc}];")
                L.append("    }")
    L.append("}")
    return name, "\n".join(L)


def emit_j3c_bf16(la, lb, lc, m0, m1):
    """Float32-tier three-centre kernel on the bf16 matrix cores for auxiliary components [m0, m1) of
    class (la lb | lc); reads the split dense pair matrix (J3cData.E32 points to it) and the vector
    kernels' items, and stores (or adds) its float32 sums into the float64 output."""
    spherical = True
    name = f"j3cb_{la}{lb}{lc}_{m0}_{m1}"
    L = la + lb + lc
    nij = n_out(la, spherical) * n_out(lb, spherical)
    nb = mfma_blocks(la, lb)
    ts = mfma_hermite(la, lb)
    nq = bf16_groups(la, lb)
    WB = bf16_dense_width(la, lb) // 8           # bf8 vectors per primitive pair
    aux_rows = aux_transform(lc, spherical)
    ms = list(range(m0, m1))
    needed = set()
    for t in ts:
        for m in ms:
            for s in aux_rows[m][1]:
                needed.add((t[0] + s[0], t[1] + s[1], t[2] + s[2]))
    P = BF16_PIECES
    prods = [(i, j) for i in range(P) for j in range(P) if i + j < P]
    out = [f"// ({la}{lb}|{lc}) aux components [{m0},{m1}) float32 tier on bf16 matrix cores ({P} pieces): L = {L}"]
    out.append("template <bool kJobs>")
    out.append(f"__device__ __forceinline__ void {name}_body(const J3cArgs a)")
    out.append("{")
    out.append("    const int item = a.item_begin + eip::xcd_remap(blockIdx.x, gridDim.x);")
    out.append("    if (item >= a.item_end) return;")
    out.append("    const eip::cptr<J3cData> jt = kJobs ? j3c_item_job(a, item) : (eip::cptr<J3cData>) nullptr;")
    out.append("#define JD(f) (kJobs ? (decltype(a.d.f)) jt->f : a.d.f)")
    out.append("    const int wave = (int) threadIdx.x >> 6, lane = (int) threadIdx.x & 63;")
    out.append("    const int kslot = lane >> 4, jcol = lane & 15;")
    out.append("    const int sp = eip::ro(a.items)[2 * item];")
    out.append("    const int ks_first = eip::ro(a.items)[2 * item + 1] + 16 * wave;")
    out.append("    const int ks_last = a.items_ks_end ? eip::ro(a.items_ks_end)[item] : (int) JD(n_ks);")
    out.append("    int* const flag = a.flags_out ? a.flags_out + (size_t) item * a.flag_stride + a.flag_slot + wave : nullptr;")
    out.append("    if (ks_first >= ks_last) {                       // no auxiliary shell for this wave")
    out.append("        if (flag && lane == 0) *flag = 0;")
    out.append("        return;")
    out.append("    }")
    out.append("    const int ks_raw = ks_first + jcol;")
    out.append("    const bool active = ks_raw < ks_last;")
    out.append("    const int ks = active ? ks_raw : ks_last - 1;")
    out.append("    const int ksh = JD(ks_perm) ? eip::ro(JD(ks_perm))[ks] : ks;")
    out.append("    const int kp0 = eip::ro(JD(ks_prim_off))[ksh], kp1 = eip::ro(JD(ks_prim_off))[ksh + 1];")
    out.append("    const int pp0 = a.items_pp ? eip::ro(a.items_pp)[2 * item] : eip::ro(JD(sp_pp_off))[sp];")
    out.append("    const int pp1 = a.items_pp ? eip::ro(a.items_pp)[2 * item + 1] : eip::ro(JD(sp_pp_off))[sp + 1];")
    out.append("    const eip::bf8* __restrict__ EB = (const eip::bf8*) JD(E32);")
    out.append("    const double* __restrict__ PP = JD(pp_P);")
    for m in ms:
        for b in range(nb):
            out.append(f"    eip::v4f acc_{m}_{b} = {{0.0f, 0.0f, 0.0f, 0.0f}};")
    # checked mixed precision: sum_k |E_k| A_{|t_k| + lc} per output row and auxiliary shell (leading
    # pieces), times the largest sum of |transform coefficients| of the chunk's components: the scale of
    # the float32 rounding of every term the row's sums contain
    for b in range(nb):
        out.append(f"    eip::v4f mag_{b} = {{0.0f, 0.0f, 0.0f, 0.0f}};")
    out.append("    const int nkp_wave = eip::wave_max(kp1 - kp0);")
    out.append("    for (int t_kp = 0; t_kp < nkp_wave; ++t_kp) {")
    out.append("      const bool kp_on = kp0 + t_kp < kp1;")
    out.append("      const int kp = kp_on ? kp0 + t_kp : kp1 - 1;")
    out.append("      const double Cx = JD(ka_C)[4 * kp + 0], Cy = JD(ka_C)[4 * kp + 1], Cz = JD(ka_C)[4 * kp + 2];")
    out.append("      const double q = JD(ka_C)[4 * kp + 3];")
    out.append("      const double ck = kp_on ? JD(ka_coef)[kp] : 0.0;")
    out.extend(_aux_prefactor_lines(lc))
    out.append("      double Pn0 = 0, Pn1 = 0, Pn2 = 0, Pn3 = 1;")
    out.append("      if (pp0 < pp1) { const int q0 = min(pp0 + kslot, pp1 - 1);")
    out.append("        Pn0 = PP[4 * (size_t) q0 + 0]; Pn1 = PP[4 * (size_t) q0 + 1]; Pn2 = PP[4 * (size_t) q0 + 2]; Pn3 = PP[4 * (size_t) q0 + 3]; }")
    out.append("      for (int g0 = pp0; g0 < pp1; g0 += 4) {")
    out.append("        const int pq = g0 + kslot;")
    out.append("        const bool pp_on = pq < pp1;")
    out.append("        const int pp = pp_on ? pq : pp1 - 1;")
    out.append("        const double Px = Pn0, Py = Pn1, Pz = Pn2, p = Pn3;")
    out.append(f"        const eip::bf8* __restrict__ Eb = EB + (size_t) pp * {WB} + jcol;")
    steps = [(qg, b) for qg in range(nq) for b in range(nb)]

    def load_step(k, indent):
        qg_, b_ = steps[k]
        return [f"{indent}const eip::bf8 ea_{k}_{pc} = Eb[{((pc * nq + qg_) * nb + b_) * 16}];" for pc in range(P)]

    # pair-matrix operands of the first steps issued before the Boys function and the Hermite
    # integrals (their latency hides behind that work); step k + BF16_E_AHEAD is issued with step k
    for k in range(min(BF16_E_AHEAD, len(steps))):
        out.extend(load_step(k, "        "))
    out.append("        { const int qn = min(g0 + 4 + kslot, pp1 - 1);      // the next group's pair (the last reloads one)")
    out.append("          Pn0 = PP[4 * (size_t) qn + 0]; Pn1 = PP[4 * (size_t) qn + 1]; Pn2 = PP[4 * (size_t) qn + 2]; Pn3 = PP[4 * (size_t) qn + 3]; }")
    # prologue in float64 (cheap), then float32 for Boys, R and the transform, as the vector float32 kernels
    out.append("        const double dPQx = Px - Cx, dPQy = Py - Cy, dPQz = Pz - Cz;")
    out.append("        double zeta, pf;")
    out.append("        if (a.omega == 0.0) {")
    out.append("            const double rs = eip::rsq_nr(p + q);")
    out.append("            zeta = p * q * (rs * rs);")
    out.append("            pf = ckq * eip::rcp_nr(p) * rs;")
    out.append("        } else {")
    out.append("            const double w2 = a.omega * a.omega;")
    out.append("            zeta = p * q * eip::rcp_nr(q + p + p * q * eip::rcp_nr(w2));")
    out.append("            pf = ckq * a.omega * eip::rcp_nr(p) * eip::rsq_nr(p * q + (p + q) * w2);")
    out.append("        }")
    out.append("        if (!pp_on) pf = 0.0;                        // an empty slot of the group adds nothing")
    out.append("        const float PQx = (float) dPQx, PQy = (float) dPQy, PQz = (float) dPQz;")
    out.append("        const float T = (float) (zeta * (dPQx * dPQx + dPQy * dPQy + dPQz * dPQz));")
    out.append(f"        float F[{L + 1}];")
    out.append(f"        eip::boys32<{L}>(T, F);")
    out.append("        const float m2z = (float) (-2.0 * zeta);")
    out.append("        const float R000_0 = (float) pf * F[0];")
    for n in range(1, L + 1):
        out.append(f"        const float pf_{n} = {'(float) pf' if n == 1 else f'pf_{n - 1}'} * m2z;")
        out.append(f"        const float R000_{n} = pf_{n} * F[{n}];")
    # Envelope of the Hermite integrals' terms: A_N^(n) = (N-1) A_{N-2}^(n+1) + |PQ| A_{N-1}^(n+1) from
    # A_0^(n) = |R000^(n)| -- the Hermite recursion with every term taken positive and |PQ| >= |PQx|,
    # |PQy|, |PQz| -- bounds the magnitude of every term summed into an R_tuv of order t + u + v = N, the
    # scale its float32 rounding (amplified by the recursion's cancellation) is relative to.
    out.append(f"        float Aenv[{L + 1}];")
    out.append("        {")
    out.append("          const float Rd = sqrtf(PQx * PQx + PQy * PQy + PQz * PQz);")
    for N in range(L + 1):
        for n in range(L - N + 1):
            if N == 0:
                ex = f"fabsf(R000_{n})"
            elif N == 1:
                ex = f"Rd * ae_0_{n + 1}"
            else:
                ex = f"{float(N - 1)}f * ae_{N - 2}_{n + 1} + Rd * ae_{N - 1}_{n + 1}"
            out.append(f"          const float ae_{N}_{n} = {ex};")
        out.append(f"          Aenv[{N}] = ae_{N}_0;")
    out.append("        }")
    rdemand = _RDemand(needed, "float") if R_ORDER == "demand" else None
    if rdemand is None:
        _emit_R(out, L, needed, "float")
    for qg in range(nq):
        out.append(f"        // Hermite indices {8 * qg} .. {min(len(ts), 8 * qg + 8) - 1}")
        for m in ms:
            r, terms = aux_rows[m]
            vals = []
            for v in range(8):
                ti = 8 * qg + v
                if ti >= len(ts):
                    vals.append("0.0f")
                    continue
                t = ts[ti]
                if rdemand is not None:
                    for s_ in terms:
                        rdemand.request(out, (t[0] + s_[0], t[1] + s_[1], t[2] + s_[2]))
                pieces = []
                for s_, (qf, j_) in terms.items():
                    assert j_ == 0
                    coef = SqrtRational(r, qf).to_float()
                    rn = _r_name(0, (t[0] + s_[0], t[1] + s_[1], t[2] + s_[2]))
                    pieces.append(rn if coef == 1.0 else f"{fmt(coef)}f * {rn}")
                vals.append("(" + " + ".join(pieces) + ")")
            out.append(f"        eip::bf8 Vb_{qg}_{m}[{P}];")
            out.append(f"        {{ const float x[8] = {{{', '.join(vals)}}};")
            out.append(f"          eip::split_bf16<{P}>(x, Vb_{qg}_{m}); }}")
        env = ", ".join(f"Aenv[{sum(ts[8 * qg + v]) + lc}]" if 8 * qg + v < len(ts) else "0.0f" for v in range(8))
        out.append(f"        eip::bf8 venv_{qg}[1];")
        out.append(f"        {{ const float x[8] = {{{env}}}; eip::split_bf16<1>(x, venv_{qg}); }}")
        for b in range(nb):
            k = steps.index((qg, b))
            if k + BF16_E_AHEAD < len(steps):
                out.extend(load_step(k + BF16_E_AHEAD, "        "))
            # products in piece order, components innermost: consecutive instructions never share an
            # accumulator
            for i_, j_ in prods:
                for m in ms:
                    out.append(f"        acc_{m}_{b} = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ea_{k}_{i_}, Vb_{qg}_{m}[{j_}], acc_{m}_{b}, 0, 0, 0);")
            out.append(f"        mag_{b} = __builtin_amdgcn_mfma_f32_16x16x32_bf16(eip::bf16_abs(ea_{k}_0), venv_{qg}[0], mag_{b}, 0, 0, 0);")
    out.append("      }")
    out.append("    }")
    cmax = max(sum(abs(SqrtRational(aux_rows[m][0], qf).to_float()) for qf, _ in aux_rows[m][1].values()) for m in ms)
    out.append("    if (flag) {                                      // this wave's largest term scale against the threshold")
    out.append(f"        const float cmax = {fmt(cmax)}f;")
    out.append("        float amax = 0.0f;")
    out.append("        if (active) {")
    for b in range(nb):
        for r_ in range(4):
            out.append(f"          if ({16 * b + r_} + 4 * kslot < {nij}) amax = fmaxf(amax, mag_{b}[{r_}]);")
    out.append("        }")
    out.append("        amax = eip::wave_max_f(cmax * amax);")
    out.append("        if (lane == 0) *flag = !(amax < (float) a.flag_threshold);   // NaN flags too")
    out.append("    }")
    out.append("    if (!active) return;")
    out.append("    const int col = eip::ro(JD(ks_col))[ksh];")
    out.append("    const int I = eip::ro(JD(sp_i))[sp], J = eip::ro(JD(sp_j))[sp];")
    out.append("    __attribute__((address_space(1))) double* __restrict__ o = (__attribute__((address_space(1))) double*) JD(out) + (JD(sp_out_base) ? eip::ro(JD(sp_out_base))[sp] : 0LL);")
    out.append("    const long long ld = JD(sp_ld) ? (long long) eip::ro(JD(sp_ld))[sp] : JD(ld);")
    out.append(f"    const long long base = (long long) sp * {nij};")
    nj = n_out(lb, spherical)
    for b in range(nb):
        for r_ in range(4):
            out.append(f"    {{ const int ij = {16 * b + r_} + 4 * kslot;")
            out.append(f"      const int mi = ij / {nj}, mj = ij - mi * {nj};")
            out.append(f"      if (ij < {nij} && (a.layout != 0 || I != J || mi >= mj)) {{")
            out.append("        const long long row = (a.layout == 0 ? eip::packed_row(I + mi, J + mj) : base + ij) * ld + col;")
            for m in ms:
                out.append(f"        if (a.accumulate) o[row + {m}] += (double) acc_{m}_{b}[{r_}]; else o[row + {m}] = (double) acc_{m}_{b}[{r_}];")
            out.append("      }")
            out.append("    }")
    out.append("#undef JD")
    out.append("}")
    out.append(f'extern "C" __global__ void __launch_bounds__({64 * MFMA_WAVES}) {name}(const J3cArgs a) {{ {name}_body<false>(a); }}')
    out.append(f'extern "C" __global__ void __launch_bounds__({64 * MFMA_WAVES}) {name}_jobs(const J3cArgs a) {{ {name}_body<true>(a); }}')
    return name, "\n".join(out)
