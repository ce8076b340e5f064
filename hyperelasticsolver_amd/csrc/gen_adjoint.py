"""Generator of the per-cell reverse-mode (vector-Jacobian product) code used by
the adjoint kernels.

The per-cell physics of the single-phase time step is written once below as
straight-line code over traced scalars: cons -> prim (HyperelasticityMPh.jl:
106-124), the Barton2009 energy split and entropy variable
(EquationsOfState.jl:118-156), the first stress row and physical flux
(EquationsOfState.jl:179-190, HyperelasticityMPh.jl:146-175) and the acoustic
tensor for n = e1 (EquationsOfState.jl:223-246, closed form).  Tracing records
an SSA list; `emit` writes, for each traced function, a primal routine and a
reverse sweep (the adjoint of every SSA statement, in reverse order) as C++
that compiles both for the GPU (hipcc) and for the host (g++, in the tests).

Run:  python hyperelasticsolver_amd/csrc/gen_adjoint.py  (writes hes_adjoint_gen.hpp)
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))


class Graph:
    def __init__(self):
        self.nodes = []  # (op, args, name)

    def add(self, op, args, name=None):
        self.nodes.append((op, tuple(args), name))
        return Var(self, len(self.nodes) - 1)

    def inp(self, name):
        return self.add("in", (), name)


def _wrap(g, x):
    if isinstance(x, Var):
        return x
    return g.add("const", (float(x),))


class Var:
    __slots__ = ("g", "i")

    def __init__(self, g, i):
        self.g, self.i = g, i

    def __add__(a, b):
        return a.g.add("add", (a, _wrap(a.g, b)))

    def __radd__(a, b):
        return a.g.add("add", (_wrap(a.g, b), a))

    def __sub__(a, b):
        return a.g.add("sub", (a, _wrap(a.g, b)))

    def __rsub__(a, b):
        return a.g.add("sub", (_wrap(a.g, b), a))

    def __mul__(a, b):
        return a.g.add("mul", (a, _wrap(a.g, b)))

    def __rmul__(a, b):
        return a.g.add("mul", (_wrap(a.g, b), a))

    def __truediv__(a, b):
        return a.g.add("div", (a, _wrap(a.g, b)))

    def __rtruediv__(a, b):
        return a.g.add("div", (_wrap(a.g, b), a))

    def __neg__(a):
        return a.g.add("neg", (a,))


def sqrt(a):
    return a.g.add("sqrt", (a,))


def sqrtp(a):
    """Square root of a positive argument, by the device's refined reciprocal
    square root (the forward step's sqrt_pos); IEEE sqrt on the host."""
    return a.g.add("sqrtp", (a,))


def rcp(a):
    """1 / a by the device's refined reciprocal (the forward step's rcp);
    IEEE division on the host."""
    return a.g.add("rcp", (a,))


def exp(a):
    return a.g.add("exp", (a,))


def clamp_min(a, c):
    """a < c ? c : a, with zero derivative on the clamped branch
    (`if S < 1e-6; S = 1e-6; end`, EquationsOfState.jl:152-154)."""
    return a.g.add("clampmin", (a, _wrap(a.g, c)))


# ----------------------------------------------------------------------------- the model
def det3(a):
    return (a[0] * (a[4] * a[8] - a[5] * a[7]) - a[3] * (a[1] * a[8] - a[7] * a[2]) +
            a[6] * (a[1] * a[5] - a[4] * a[2]))


def constants(b0, c0, rho0, cv, t0):
    """Barton2009 derived constants with alpha = 1, beta = 3, gamma = 2
    (EquationsOfState.jl:90, 111-112)."""
    b0sq = b0 * b0
    k0 = c0 * c0 - (4.0 / 3.0) * b0sq
    return dict(rho0=rho0, cv=cv, t0=t0, b0sq=b0sq, k0=k0, hk0a2=0.5 * k0, k0_2a=0.5 * k0,
                k0_2=0.5 * k0, cvt0=cv * t0, hb0sq=0.5 * b0sq)


def thermo(Q, m):
    """Q: [rho, rho u(3), rho E, rho A(9)] -> derived state of the phase, in the
    form the forward step uses: G = (A A^T)^-1 from P = rho A as
    adj(P P^T) / (rho rho0)^2, I3 = (rho / rho0)^2."""
    P = Q[5:14]
    detP = det3(P)
    irho0 = rcp(m["rho0"])
    rho = sqrtp(detP * irho0)
    iden = rcp(rho)
    u = [Q[1] * iden, Q[2] * iden, Q[3] * iden]
    E = Q[4] * iden
    e = E - 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2])
    B11 = P[0] * P[0] + P[3] * P[3] + P[6] * P[6]
    B22 = P[1] * P[1] + P[4] * P[4] + P[7] * P[7]
    B33 = P[2] * P[2] + P[5] * P[5] + P[8] * P[8]
    B12 = P[0] * P[1] + P[3] * P[4] + P[6] * P[7]
    B13 = P[0] * P[2] + P[3] * P[5] + P[6] * P[8]
    B23 = P[1] * P[2] + P[4] * P[5] + P[7] * P[8]
    ir = iden * irho0
    sc = ir * ir
    G = dict(G11=(B22 * B33 - B23 * B23) * sc, G22=(B11 * B33 - B13 * B13) * sc,
             G33=(B11 * B22 - B12 * B12) * sc, G12=(B13 * B23 - B12 * B33) * sc,
             G13=(B12 * B23 - B13 * B22) * sc, G23=(B12 * B13 - B11 * B23) * sc)
    i1 = G["G11"] + G["G22"] + G["G33"]
    trG2 = (G["G11"] * G["G11"] + G["G22"] * G["G22"] + G["G33"] * G["G33"] +
            2.0 * (G["G12"] * G["G12"] + G["G13"] * G["G13"] + G["G23"] * G["G23"]))
    i2 = 0.5 * (i1 * i1 - trG2)
    s1 = rho * irho0  # I3^(1/2) = 1/det A = rho/rho0
    i3 = s1 * s1
    sa = s1
    sb = i3 * s1
    sg = i3
    r0i = m["rho0"] * iden
    ith = (r0i * r0i) * rcp(m["cvt0"])
    W = m["hb0sq"] * sb * (i1 * i1 * (1.0 / 3.0) - i2)
    sam1 = sa - 1.0
    Uc = m["hk0a2"] * sam1 * sam1
    x = clamp_min((e - W - Uc) * ith + 1.0, 1e-6)
    Eth = m["cvt0"] * sg * (x - 1.0)
    t = dict(rho=rho, u=u, E=E, P=P, i1=i1, i2=i2, i3=i3, sa=sa, sb=sb, sg=sg, W=W, x=x, Eth=Eth)
    t.update(G)
    return t


def c3_of(m, t):
    return m["k0_2a"] * (t["sa"] - 1.0) * t["sa"] + 1.0 * t["Eth"] + 1.5 * t["W"]


def flux(m, t):
    G11, G22, G33, G12, G13, G23 = (t[k] for k in ("G11", "G22", "G33", "G12", "G13", "G23"))
    G2_11 = G11 * G11 + G12 * G12 + G13 * G13
    G2_12 = G11 * G12 + G12 * G22 + G13 * G23
    G2_13 = G11 * G13 + G12 * G23 + G13 * G33
    kap = m["hb0sq"] * t["sb"]
    c3 = c3_of(m, t)
    i13 = t["i1"] * (1.0 / 3.0)
    m2r = -2.0 * t["rho"]
    s0 = m2r * (kap * (G2_11 - i13 * G11) + c3)
    s1 = m2r * (kap * (G2_12 - i13 * G12))
    s2 = m2r * (kap * (G2_13 - i13 * G13))
    u, P, den = t["u"], t["P"], t["rho"]
    du1 = den * u[0]
    F = [du1, du1 * u[0] - s0, du1 * u[1] - s1, du1 * u[2] - s2,
         du1 * t["E"] - (u[0] * s0 + u[1] * s1 + u[2] * s2)]
    FA = [None] * 9
    for c in range(3):
        FA[3 * c] = None  # row 1 of the distortion flux is identically zero
        # rho (u1 A_kc - u_k A_1c) with rho A = P
        FA[3 * c + 1] = u[0] * P[1 + 3 * c] - u[1] * P[3 * c]
        FA[3 * c + 2] = u[0] * P[2 + 3 * c] - u[2] * P[3 * c]
    return F + FA


def acoustic_e1(m, t):
    """O11 O12 O13 O22 O23 O33 for n = e1 in the collected form the forward
    step uses (alpha = 1, beta = 3, gamma = 2): with T = G^2 - (I1/3) G and
    2 kappa = b0^2 I3^(3/2),
      O11 = 2 kappa (9 T11 + 4/3 G11^2) + s_a k0 (3 s_a - 2) + 6 E_th + 12 W,
      O1k = 2 kappa (5 T1k + 4/3 G11 G1k),  Ojk = 2 kappa (Tjk + G11 Gjk + 1/3 G1j G1k)."""
    G11, G22, G33, G12, G13, G23 = (t[k] for k in ("G11", "G22", "G33", "G12", "G13", "G23"))
    i13 = t["i1"] * (1.0 / 3.0)
    T11 = G11 * G11 + G12 * G12 + G13 * G13 - i13 * G11
    T12 = G11 * G12 + G12 * G22 + G13 * G23 - i13 * G12
    T13 = G11 * G13 + G12 * G23 + G13 * G33 - i13 * G13
    T22 = G12 * G12 + G22 * G22 + G23 * G23 - i13 * G22
    T23 = G12 * G13 + G22 * G23 + G23 * G33 - i13 * G23
    T33 = G13 * G13 + G23 * G23 + G33 * G33 - i13 * G33
    k2 = m["b0sq"] * t["sb"]
    g43 = (4.0 / 3.0) * G11
    sa = t["sa"]
    O11 = k2 * (9.0 * T11 + g43 * G11) + (m["k0"] * sa * (3.0 * sa - 2.0) + 6.0 * t["Eth"] +
                                          12.0 * t["W"])
    O12 = k2 * (5.0 * T12 + g43 * G12)
    O13 = k2 * (5.0 * T13 + g43 * G13)
    g12 = (1.0 / 3.0) * G12
    g13 = (1.0 / 3.0) * G13
    O22 = k2 * (T22 + G11 * G22 + g12 * G12)
    O23 = k2 * (T23 + G11 * G23 + g12 * G13)
    O33 = k2 * (T33 + G11 * G33 + g13 * G13)
    return [O11, O12, O13, O22, O23, O33]


def energy_from_prims(m, A, S):
    """e(S, G(A)) for prim -> cons (EquationsOfState.jl:118-137), det A > 0."""
    B11 = A[0] * A[0] + A[3] * A[3] + A[6] * A[6]
    B22 = A[1] * A[1] + A[4] * A[4] + A[7] * A[7]
    B33 = A[2] * A[2] + A[5] * A[5] + A[8] * A[8]
    B12 = A[0] * A[1] + A[3] * A[4] + A[6] * A[7]
    B13 = A[0] * A[2] + A[3] * A[5] + A[6] * A[8]
    B23 = A[1] * A[2] + A[4] * A[5] + A[7] * A[8]
    C11 = B22 * B33 - B23 * B23
    C12 = B13 * B23 - B12 * B33
    C13 = B12 * B23 - B13 * B22
    detB = B11 * C11 + B12 * C12 + B13 * C13
    idetB = 1.0 / detB
    G11, G22, G33 = C11 * idetB, (B11 * B33 - B13 * B13) * idetB, (B11 * B22 - B12 * B12) * idetB
    G12, G13, G23 = C12 * idetB, C13 * idetB, (B12 * B13 - B11 * B23) * idetB
    i1 = G11 + G22 + G33
    trG2 = G11 * G11 + G22 * G22 + G33 * G33 + 2.0 * (G12 * G12 + G13 * G13 + G23 * G23)
    i2 = 0.5 * (i1 * i1 - trG2)
    i3 = idetB
    sa = 1.0 / det3(A)
    sb = i3 * sa
    sg = i3
    sam1 = sa - 1.0
    return (m["hk0a2"] * sam1 * sam1 + m["cvt0"] * sg * (exp(S * (1.0 / m["cv"])) - 1.0) +
            m["hb0sq"] * sb * (i1 * i1 * (1.0 / 3.0) - i2))


# ----------------------------------------------------------------------------- traced functions
def trace_cell():
    """in: Q(14), b0, c0, rho0, cv, t0 -> out: F(14), Omega(6: O11 O12 O13 O22 O23 O33), u1"""
    g = Graph()
    Q = [g.inp(f"q{k}") for k in range(14)]
    b0, c0, rho0, cv, t0 = (g.inp(n) for n in ("b0", "c0", "rho0", "cv", "t0"))
    m = constants(b0, c0, rho0, cv, t0)
    t = thermo(Q, m)
    F = flux(m, t)
    O = acoustic_e1(m, t)
    inputs = Q + [b0, c0, rho0, cv, t0]
    return g, inputs, F + O + [t["u"][0]]


def trace_prim2cons():
    """in: u(3), A(9), S, b0, c0, rho0, cv, t0 -> out: Q(14) (HyperelasticityMPh.jl:66-87, a = 1)."""
    g = Graph()
    u = [g.inp(f"u{k}") for k in range(3)]
    A = [g.inp(f"a{k}") for k in range(9)]
    S = g.inp("S")
    b0, c0, rho0, cv, t0 = (g.inp(n) for n in ("b0", "c0", "rho0", "cv", "t0"))
    m = constants(b0, c0, rho0, cv, t0)
    rho = rho0 / det3(A)
    e = energy_from_prims(m, A, S)
    ek = (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / 2.0
    Q = [rho, rho * u[0], rho * u[1], rho * u[2], rho * (e + ek)] + [rho * a for a in A]
    return g, u + A + [S, b0, c0, rho0, cv, t0], Q


def observed(Q, rho0):
    """rho and u1 of a cell as cons2prim gives them (HyperelasticityMPh.jl:113-116)."""
    rho = sqrt(det3(Q[5:14]) * (1.0 / rho0))
    return rho, Q[1] * (1.0 / rho)


def trace_loss():
    """in: Q(14), rho*, u1*, rho0 -> out: (rho - rho*)^2 + (u1 - u1*)^2 with rho, u1 from cons2prim."""
    g = Graph()
    Q = [g.inp(f"q{k}") for k in range(14)]
    rt, ut, rho0 = g.inp("rho_t"), g.inp("u_t"), g.inp("rho0")
    rho, u1 = observed(Q, rho0)
    d1 = rho - rt
    d2 = u1 - ut
    return g, Q + [rt, ut, rho0], [d1 * d1 + d2 * d2]


def trace_profiles():
    """in: Q(14), rho0 -> out: rho, u1 (the observed quantities of the misfit)."""
    g = Graph()
    Q = [g.inp(f"q{k}") for k in range(14)]
    rho0 = g.inp("rho0")
    rho, u1 = observed(Q, rho0)
    return g, Q + [rho0], [rho, u1]


# ----------------------------------------------------------------------------- emission
def _lit(x):
    r = repr(float(x))
    return r if ("." in r or "e" in r or "inf" in r or "nan" in r) else r + ".0"


def _ancestors(g, roots):
    seen = set()
    stack = [o.i for o in roots if o is not None]
    while stack:
        n = stack.pop()
        if n in seen:
            continue
        seen.add(n)
        stack.extend(a.i for a in g.nodes[n][1] if isinstance(a, Var))
    return seen


def emit(fname, g, inputs, outputs, fused_split=None, exact=False):
    """C++ for `fname_fwd(in, out)` and `fname_vjp(in, obar, ibar)` (ibar accumulates).

    With fused_split = k, also `fname_fwd_vjp(in, out, seed, ibar)`: the primal,
    then obar from `seed(out, obar)` (which returns whether obar[k:] may be
    non-zero), then the reverse sweep.  The nodes that only feed outputs k..
    are swept first, inside `if`, so a zero tail of obar skips them: none of
    them is an ancestor of a node that feeds outputs 0..k-1, so their
    adjoints are complete before the shared nodes are reached."""
    idx_in = {v.i: k for k, v in enumerate(inputs)}
    live = _ancestors(g, outputs)  # nodes reachable from outputs
    order = sorted(live)

    def val(a):
        op, args, _ = g.nodes[a.i]
        return _lit(args[0]) if op == "const" else f"v{a.i}"

    fwd = []
    for n in order:
        op, args, _ = g.nodes[n]
        if op == "in":
            fwd.append(f"  const double v{n} = in[{idx_in[n]}];")
        elif op == "const":
            continue
        elif op in ("add", "sub", "mul", "div"):
            sym = {"add": "+", "sub": "-", "mul": "*", "div": "/"}[op]
            fwd.append(f"  const double v{n} = {val(args[0])} {sym} {val(args[1])};")
        elif op == "neg":
            fwd.append(f"  const double v{n} = -{val(args[0])};")
        elif op == "sqrt":
            fwd.append(f"  const double v{n} = sqrt({val(args[0])});")
        elif op == "sqrtp":
            fwd.append(f"  const double v{n} = gen_sqrt({val(args[0])});")
        elif op == "rcp":
            fwd.append(f"  const double v{n} = gen_rcp({val(args[0])});")
        elif op == "exp":
            fwd.append(f"  const double v{n} = exp({val(args[0])});")
        elif op == "clampmin":
            fwd.append(f"  const double v{n} = {val(args[0])} < {val(args[1])} ? {val(args[1])} : {val(args[0])};")
        else:
            raise ValueError(op)
    nout = len(outputs)
    outs_fwd = [f"  out[{k}] = {val(o) if o is not None else '0.0'};" for k, o in enumerate(outputs)]

    def is_var(a):
        return g.nodes[a.i][0] != "const"

    def rev_node(n):
        op, args, _ = g.nodes[n]
        if op in ("in", "const"):
            return []
        gn = f"g{n}"
        a0 = args[0]
        a1 = args[1] if len(args) > 1 else None
        r = []
        if op == "add":
            if is_var(a0): r.append(f"  g{a0.i} += {gn};")
            if is_var(a1): r.append(f"  g{a1.i} += {gn};")
        elif op == "sub":
            if is_var(a0): r.append(f"  g{a0.i} += {gn};")
            if is_var(a1): r.append(f"  g{a1.i} -= {gn};")
        elif op == "mul":
            if is_var(a0): r.append(f"  g{a0.i} += {gn} * {val(a1)};")
            if is_var(a1): r.append(f"  g{a1.i} += {gn} * {val(a0)};")
        elif op == "div":
            if is_var(a0): r.append(f"  g{a0.i} += {gn} / {val(a1)};")
            if is_var(a1): r.append(f"  g{a1.i} -= {gn} * v{n} / {val(a1)};")
        elif op == "neg":
            r.append(f"  g{a0.i} -= {gn};")
        elif op == "sqrt":
            r.append(f"  g{a0.i} += {gn} * (0.5 / v{n});")
        elif op == "sqrtp":
            r.append(f"  g{a0.i} += {gn} * (0.5 * gen_rcp(v{n}));")
        elif op == "rcp":
            if is_var(a0): r.append(f"  g{a0.i} -= {gn} * (v{n} * v{n});")
        elif op == "exp":
            r.append(f"  g{a0.i} += {gn} * v{n};")
        elif op == "clampmin":
            r.append(f"  if (!({val(a0)} < {val(a1)})) g{a0.i} += {gn};")
        return r

    init = [f"  double g{n} = 0.0;" for n in order if g.nodes[n][0] != "const"]
    seeds = [f"  g{o.i} += obar[{k}];" for k, o in enumerate(outputs)
             if o is not None and g.nodes[o.i][0] != "const"]
    rev = []
    for n in reversed(order):
        rev += rev_node(n)
    rev_out = [f"  ibar[{k}] += g{v.i};" for k, v in enumerate(inputs) if v.i in live]
    nin = len(inputs)
    head = f"// {fname}: {nin} inputs, {nout} outputs, {len(order)} SSA values\n"
    # exact: no contraction into fused multiply-adds, so every kernel that
    # inlines the routine rounds it the same way
    pre = "#pragma clang fp contract(off)\n" if exact else ""
    f1 = (f"HES_HD inline void {fname}_fwd(const double* in, double* out) {{\n" + pre + "\n".join(fwd) + "\n" +
          "\n".join(outs_fwd) + "\n}\n")
    f2 = (f"HES_HD inline void {fname}_vjp(const double* in, const double* obar, double* ibar) {{\n" + pre +
          "\n".join(fwd) + "\n" + "\n".join(init + seeds + rev) + "\n" + "\n".join(rev_out) + "\n}\n")
    text = head + f1 + "\n" + f2
    if fused_split is not None:
        # fwd_vjp: the primal of the nodes that feed outputs 0..k-1 (and the
        # outputs among them), `seed_head(out, obar)` for those outputs, which
        # returns whether the remaining outputs' cotangents may be non-zero;
        # only then the primal of the nodes that feed only the remaining
        # outputs, `seed_tail(out, obar)` for them and their reverse sweep;
        # last the reverse sweep of the first nodes.  No node of the second
        # group is an ancestor of the first, so its adjoints are complete
        # before the first group's sweep.
        head_set = _ancestors(g, outputs[:fused_split])
        nodes_fwd = [n for n in order if g.nodes[n][0] != "const"]
        fwd_head = [l for l, n in zip(fwd, nodes_fwd) if n in head_set]
        fwd_tail = [l for l, n in zip(fwd, nodes_fwd) if n not in head_set]
        head_out = [k for k, o in enumerate(outputs) if o is None or o.i in head_set]
        tail_out = [k for k in range(nout) if k not in head_out]
        tail_only = [n for n in reversed(order) if n not in head_set]
        rest = [n for n in reversed(order) if n in head_set]
        rev_tail = [r for n in tail_only for r in rev_node(n)]
        rev_rest = [r for n in rest for r in rev_node(n)]
        seed_of = lambda ks: [f"  g{outputs[k].i} += obar[{k}];" for k in ks
                              if outputs[k] is not None and g.nodes[outputs[k].i][0] != "const"]
        f3 = (f"template <class SeedHead, class SeedTail>\n"
              f"HES_HD inline void {fname}_fwd_vjp(const double* in, double* out, SeedHead&& seed_head, "
              f"SeedTail&& seed_tail, double* ibar) {{\n" + "\n".join(fwd_head) + "\n" +
              "\n".join(outs_fwd[k] for k in head_out) + "\n"
              f"  double obar[{nout}];\n"
              f"  const bool tail = seed_head(static_cast<const double*>(out), obar);\n" +
              "\n".join(init + seed_of(head_out)) + "\n  if (tail) {\n" +
              "\n".join("  " + l for l in fwd_tail) + "\n" +
              "\n".join("  " + outs_fwd[k] for k in tail_out) + "\n"
              f"    seed_tail(static_cast<const double*>(out), obar);\n" +
              "\n".join("  " + l for l in seed_of(tail_out)) + "\n" +
              "\n".join("  " + r for r in rev_tail) +
              "\n  }\n" + "\n".join(rev_rest) + "\n" + "\n".join(rev_out) + "\n}\n")
        text += "\n" + f3
    return text


def main(path=os.path.join(HERE, "hes_adjoint_gen.hpp")):
    parts = [
        "// hes_adjoint_gen.hpp -- GENERATED by gen_adjoint.py; do not edit.\n"
        "// Primal and reverse-mode (VJP) routines of the single-phase per-cell physics.\n"
        "#pragma once\n#include <cmath>\n"
        "#if !defined(__clang__)\n#pragma GCC diagnostic ignored \"-Wunknown-pragmas\"\n#endif\n"
        "#if defined(__HIPCC__)\n#define HES_HD __host__ __device__\n#else\n#define HES_HD\n#endif\n"
        "namespace hes_gen {\nusing std::sqrt;\nusing std::exp;\n"
        "// 1/x and sqrt(x) as the forward step forms them on the device (refined\n"
        "// hardware reciprocal and reciprocal square root); IEEE on the host\n"
        "HES_HD inline double gen_rcp(double x) {\n"
        "#if defined(__HIP_DEVICE_COMPILE__)\n"
        "  double r = __builtin_amdgcn_rcp(x);\n"
        "  r = fma(fma(-x, r, 1.0), r, r);\n"
        "  return fma(fma(-x, r, 1.0), r, r);\n"
        "#else\n  return 1.0 / x;\n#endif\n}\n"
        "HES_HD inline double gen_sqrt(double x) {\n"
        "#if defined(__HIP_DEVICE_COMPILE__)\n"
        "  const double y = __builtin_amdgcn_rsq(x);\n"
        "  double g = x * y, h = 0.5 * y;\n"
        "  const double e = fma(-h, g, 0.5);\n"
        "  g = fma(g, e, g);\n  h = fma(h, e, h);\n"
        "  return fma(fma(-g, g, x), h, g);\n"
        "#else\n  return sqrt(x);\n#endif\n}\n"
        "constexpr int CELL_IN = 19, CELL_OUT = 21, P2C_IN = 18, P2C_OUT = 14, LOSS_IN = 17, PROF_IN = 15;\n"]
    g, i, o = trace_cell()
    parts.append(emit("cell", g, i, o, fused_split=14))
    g, i, o = trace_prim2cons()
    parts.append(emit("prim2cons", g, i, o))
    g, i, o = trace_loss()
    parts.append(emit("loss", g, i, o))
    g, i, o = trace_profiles()
    parts.append(emit("profiles", g, i, o, exact=True))
    parts.append("}  // namespace hes_gen\n")
    text = "\n".join(parts)
    with open(path, "w") as f:
        f.write(text)
    return path


if __name__ == "__main__":
    print(main())
