"""Independent NumPy FP64 restatement of the HOME-LBM step.

TEST INFRASTRUCTURE ONLY (imported by tests/). It is written from the
tensor form of the method rather than from the reference's unrolled sums, so
an index, sign or weight slip in either this file or the C oracle shows up
as a disagreement between the two:

* D2Q9 reconstruction (envs/lbm/lbm_func.py:685-788) as the third-order
  Hermite expansion  f_i = w_i rho [1 + 3 c.u + 4.5 Q_i:S
  + 13.5 (H_xxy,i A_xxy + H_xyy,i A_xyy)],  Q = cc - I/3,
  H_xxy = cx^2 cy - cy/3,  H_xyy = cx cy^2 - cx/3.
* D3Q27 reconstruction (envs/lbm3d/lbm_func_3d.py:224-251) as
  f_i = w_i rho [1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u + 1.5 Q_i:S].
* Collision (lbm_func.py:433-461, lbm_func_3d.py:579-595) as trace /
  deviator relaxation: the trace of Pi is set to its equilibrium value, the
  deviator relaxes with rate Omega towards dev(rho u u), plus the forcing
  terms.
* Boundary faces (lbm_func.py:634-682, lbm_func_3d.py:627-686) in the order
  x, then y, then z.
* Wall rule on mask-cut links at the link midpoint (lbm_func.py:349-401,
  lbm_func_3d.py:446-511).

Arrays use the oracle layout: (nfields, ny, nx) or (nfields, nz, ny, nx).
"""
from __future__ import annotations

import numpy as np

D2_C = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])
D2_W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
D3_C = np.array([
    [0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1],
    [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1],
    [1, 1, 1], [-1, -1, -1], [1, 1, -1], [-1, -1, 1], [1, -1, 1], [-1, 1, -1],
    [-1, 1, 1], [1, -1, -1]])
D3_W = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)


def inverse_index(c: np.ndarray) -> np.ndarray:
    """Index of -c_i for every i, found by search (not tabulated)."""
    return np.array([int(np.where((c == -ci).all(axis=1))[0][0]) for ci in c])


# ---------------------------------------------------------------- tensors

def _unpack(m: np.ndarray, lattice: int):
    """Moments -> (rho, u[d], S[d, d]) with S symmetric."""
    if lattice == 9:
        rho, ux, uy, sxx, syy, sxy = m
        u = np.stack([ux, uy])
        S = np.stack([np.stack([sxx, sxy]), np.stack([sxy, syy])])
    else:
        rho, ux, uy, uz, sxx, syy, szz, sxy, sxz, syz = m
        u = np.stack([ux, uy, uz])
        S = np.stack([np.stack([sxx, sxy, sxz]), np.stack([sxy, syy, syz]), np.stack([sxz, syz, szz])])
    return rho, u, S


def _pack(rho, u, S, lattice: int) -> np.ndarray:
    if lattice == 9:
        return np.stack([rho, u[0], u[1], S[0, 0], S[1, 1], S[0, 1]])
    return np.stack([rho, u[0], u[1], u[2], S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])


def recon(lattice: int, rho, u, S, i: int):
    """Population i rebuilt from (rho, u, S); broadcasts over cells."""
    if lattice == 9:
        c, w, d = D2_C[i].astype(float), D2_W[i], 2
    else:
        c, w, d = D3_C[i].astype(float), D3_W[i], 3
    cu = sum(c[a] * u[a] for a in range(d))
    QS = sum((c[a] * c[b] - (a == b) / 3.0) * S[a, b] for a in range(d) for b in range(d))
    if lattice == 9:
        ux, uy = u
        a_xxy = S[0, 0] * uy + 2 * S[0, 1] * ux - 2 * ux * ux * uy
        a_xyy = S[1, 1] * ux + 2 * S[0, 1] * uy - 2 * ux * uy * uy
        h_xxy = c[0] ** 2 * c[1] - c[1] / 3.0
        h_xyy = c[0] * c[1] ** 2 - c[0] / 3.0
        return w * rho * (1 + 3 * cu + 4.5 * QS + 13.5 * (h_xxy * a_xxy + h_xyy * a_xyy))
    uu = sum(u[a] * u[a] for a in range(3))
    return w * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * uu + 1.5 * QS)


def s_equilibrium(lattice: int, rho, u, clamp: bool):
    """S of the plain second-order equilibrium, from its populations."""
    C = D2_C if lattice == 9 else D3_C
    W = D2_W if lattice == 9 else D3_W
    d = C.shape[1]
    uu = sum(u[a] * u[a] for a in range(d))
    Pi = np.zeros((d, d) + np.shape(rho))
    for i in range(len(W)):
        cu = sum(C[i, a] * u[a] for a in range(d))
        f = W[i] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * uu)
        if clamp:
            f = np.maximum(f, 0.0)
        Pi += np.einsum("a,b->ab", C[i], C[i])[(...,) + (None,) * np.ndim(rho)] * f
    with np.errstate(invalid="ignore", divide="ignore"):  # unset edge cells are overwritten later
        return Pi / rho - np.eye(d)[(...,) + (None,) * np.ndim(rho)] / 3.0


def collide(lattice: int, rho, j, Pi, F, omega):
    """Trace/deviator form of the regularized collision with forcing.

    Returns (u_store, S_store) as the reference stores them."""
    d = 2 if lattice == 9 else 3
    eye = np.eye(d)[(...,) + (None,) * np.ndim(rho)]
    u = (j + 0.5 * F) / rho
    ruu = rho * np.einsum("a...,b...->ab...", u, u)
    G = 0.5 * (np.einsum("a...,b...->ab...", F, u) + np.einsum("a...,b...->ab...", u, F))
    if d == 2:
        def dev(A):  # 2D deviator
            t = 0.5 * (A[0, 0] + A[1, 1])
            return A - t * eye
        post = (rho / 3.0) * eye + 0.5 * (ruu[0, 0] + ruu[1, 1]) * eye \
            + (1 - omega) * dev(Pi) + omega * dev(ruu) \
            + 0.5 * (G[0, 0] + G[1, 1]) * eye + (2 - omega) * dev(G)
    else:
        def dev(A):
            t = (A[0, 0] + A[1, 1] + A[2, 2]) / 3.0
            return A - t * eye
        post = (rho / 3.0) * eye + (ruu[0, 0] + ruu[1, 1] + ruu[2, 2]) / 3.0 * eye \
            + (1 - omega) * dev(Pi) + omega * dev(ruu) + G
    return u + F / (2 * rho), post / rho - eye / 3.0


# ------------------------------------------------------------------- step

def _wall_velocity(lattice: int, solids, sid: np.ndarray, h: np.ndarray):
    """u_p at points h[d, n] of solids sid[n] (pose tables as in the C oracle)."""
    d = 2 if lattice == 9 else 3
    up = np.zeros((d, h.shape[1]))
    r = np.zeros((d, h.shape[1]))
    if solids is None:
        return up, r
    solids = np.asarray(solids, dtype=np.float64)
    for k in np.unique(sid):
        sel = sid == k
        if k < 0 or k >= solids.shape[0]:
            continue
        P = solids[k]
        if d == 2:
            p, ang, pl, angl = P[0:2], P[2:4], P[4:6], P[6:8]
            Rn = np.array([[ang[0], -ang[1]], [ang[1], ang[0]]])
            Rl = np.array([[angl[0], -angl[1]], [angl[1], angl[0]]])
        else:
            p, Rn, pl, Rl = P[0:3], P[3:12].reshape(3, 3), P[12:15], P[15:24].reshape(3, 3)
        rel = h[:, sel] - p[:, None]
        q = Rn.T @ rel
        up[:, sel] = (Rn @ q + p[:, None]) - (Rl @ q + pl[:, None])
        r[:, sel] = rel
    return up, r


def step(lattice: int, m: np.ndarray, nu: float, bc_type, bc_value, mask=None, force=None,
         solids=None) -> np.ndarray:
    """One FP64 time step; returns the new moment array."""
    m = np.asarray(m, dtype=np.float64)
    d = 2 if lattice == 9 else 3
    C = D2_C if lattice == 9 else D3_C
    shape = m.shape[1:]
    inner = tuple(slice(1, n - 1) for n in shape)
    rho, u, S = _unpack(m, lattice)
    rho_c, u_c, S_c = rho[inner], u[(slice(None),) + inner], S[(slice(None), slice(None)) + inner]
    npop = len(C)
    pops = []
    if mask is not None:
        occ = np.asarray(mask) != 0
        occ_c = occ[inner]
        mval = np.asarray(mask).astype(np.int64)
    # axis order of arrays is (z,) y, x; c is (x, y[, z])
    coords = np.meshgrid(*[np.arange(1, n - 1) for n in shape], indexing="ij")
    pos = np.stack(coords[::-1]).astype(float)  # (x, y[, z]) of inner cells
    for i in range(npop):
        ci = C[i][::-1]  # in array-axis order
        src = tuple(slice(1 - ci[a], shape[a] - 1 - ci[a]) for a in range(d))
        f = recon(lattice, rho[src], u[(slice(None),) + src], S[(slice(None), slice(None)) + src], i)
        f = np.maximum(f, 0.0)
        if mask is not None:
            cut = occ[src] != occ_c
            if cut.any():
                sid = np.where(occ_c, mval[inner] - 1, mval[src] - 1)[cut]
                h = pos[:, cut] - 0.5 * C[i][:, None]
                up, _ = _wall_velocity(lattice, solids, sid, h)
                uc = u_c[:, cut]
                Ss = np.einsum("a...,b...->ab...", up, up) + S_c[:, :, cut] \
                    - np.einsum("a...,b...->ab...", uc, uc)
                fw = recon(lattice, rho_c[cut], up, Ss, i)
                if lattice == 9:
                    fw = np.maximum(fw, 0.0)
                f = f.copy()
                f[cut] = fw
        pops.append(f)
    pops = np.array(pops)
    rho_n = pops.sum(axis=0)
    j = np.einsum("ia,i...->a...", C.astype(float), pops)
    Pi = np.einsum("ia,ib,i...->ab...", C.astype(float), C.astype(float), pops)
    if force is None:
        F = np.zeros((d,) + rho_n.shape)
    else:
        F = np.asarray(force, dtype=np.float64)[(slice(None),) + inner]
    omega = 1.0 / (3.0 * nu + 0.5)
    u_n, S_n = collide(lattice, rho_n, j, Pi, F, omega)
    R = np.zeros(shape)
    U = np.zeros((d,) + shape)
    SS = np.zeros((d, d) + shape)
    R[inner] = rho_n
    U[(slice(None),) + inner] = u_n
    SS[(slice(None), slice(None)) + inner] = S_n
    _faces(lattice, R, U, SS, bc_type, bc_value)
    return _pack(R, U, SS, lattice)


def _faces(lattice: int, R, U, SS, bc_type, bc_value):
    """Face passes x, y, z. Face ids: 2D (left 0, top 1, right 2, bottom 3);
    3D (left 0, right 1, top 2, bottom 3, front 4, back 5)."""
    d = 2 if lattice == 9 else 3
    if d == 2:
        faces = [(0, 0, 0), (0, -1, 2), (1, -1, 1), (1, 0, 3)]
    else:
        faces = [(0, 0, 0), (0, -1, 1), (1, -1, 2), (1, 0, 3), (2, -1, 4), (2, 0, 5)]
    for axis, end, fid in faces:
        ax = d - 1 - axis  # array axis of this spatial axis
        b = [slice(None)] * d
        n = [slice(None)] * d
        b[ax] = end
        n[ax] = 1 if end == 0 else -2
        b, n = tuple(b), tuple(n)
        rho_n = R[n]
        u_n = U[(slice(None),) + n]
        S_nb = SS[(slice(None), slice(None)) + n]
        t = int(bc_type[fid])
        if t == 0:
            u_b = np.broadcast_to(np.asarray(bc_value[fid][:d], dtype=float)[(...,) + (None,) * (d - 1)],
                                  u_n.shape).copy()
        elif t == 1:
            u_b = u_n.copy()
        elif t == 2:
            u_b = u_n.copy()
            u_b[axis] = 0.0
        else:
            u_b = U[(slice(None),) + b].copy()
        clamp = lattice == 9
        S_b = s_equilibrium(lattice, rho_n, u_b, clamp) - s_equilibrium(lattice, rho_n, u_n, clamp) + S_nb
        R[b] = rho_n
        U[(slice(None),) + b] = u_b
        SS[(slice(None), slice(None)) + b] = S_b


def forces(lattice: int, m: np.ndarray, mask: np.ndarray, solids) -> np.ndarray:
    """Momentum-exchange force/torque per solid, FP64: [nsolid, 6]."""
    m = np.asarray(m, dtype=np.float64)
    d = 2 if lattice == 9 else 3
    C = D2_C if lattice == 9 else D3_C
    inv = inverse_index(C)
    shape = m.shape[1:]
    nsolid = 0 if solids is None else np.asarray(solids).shape[0]
    out = np.zeros((nsolid, 6))
    inner = tuple(slice(1, n - 1) for n in shape)
    rho, u, S = _unpack(m, lattice)
    occ = np.asarray(mask) != 0
    mval = np.asarray(mask).astype(np.int64)
    coords = np.meshgrid(*[np.arange(1, n - 1) for n in shape], indexing="ij")
    pos = np.stack(coords[::-1]).astype(float)
    nx = shape[-1]
    for i in range(len(C)):
        ci = C[i][::-1]
        src = tuple(slice(1 - ci[a], shape[a] - 1 - ci[a]) for a in range(d))
        cut = occ[src] != occ[inner]
        if not cut.any():
            continue
        sid = np.where(occ[inner], mval[inner] - 1, mval[src] - 1)[cut]
        h = pos[:, cut] - 0.5 * C[i][:, None]
        up, r = _wall_velocity(lattice, solids, sid, h)
        rc, uc, Sc = rho[inner][cut], u[(slice(None),) + inner][:, cut], S[(slice(None), slice(None)) + inner][:, :, cut]
        Ss = np.einsum("a...,b...->ab...", up, up) + Sc - np.einsum("a...,b...->ab...", uc, uc)
        fw = recon(lattice, rc, up, Ss, i)
        if lattice == 9:
            fw = np.maximum(fw, 0.0)
        fcur = recon(lattice, rc, uc, Sc, inv[i])
        Fv = fcur * (C[inv[i]][:, None] - up) - fw * (C[i][:, None] - up)
        for k in np.unique(sid):
            if k < 0 or k >= nsolid:
                continue
            sel = sid == k
            out[k, :d] += Fv[:, sel].sum(axis=1)
            if d == 2:
                out[k, 5] += ((r[0, sel] * Fv[1, sel] - r[1, sel] * Fv[0, sel]) / nx).sum()
            else:
                out[k, 3:6] += np.cross(r[:, sel].T, Fv[:, sel].T).sum(axis=0)
    return out


def voxelize_f32(shape, capsules: np.ndarray) -> np.ndarray:
    """Vectorised FP32 capsule voxelizer, same evaluation order as the C one."""
    caps = np.asarray(capsules, dtype=np.float32).reshape(-1, 8)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float32) for n in shape], indexing="ij")
    pts = list(grids[::-1]) + [np.zeros(shape, np.float32)] * (3 - len(shape))
    px, py, pz = pts[0], pts[1], pts[2]
    out = np.zeros(shape, dtype=np.uint8)
    done = np.zeros(shape, dtype=bool)
    for k in range(caps.shape[0]):
        a = caps[k, 0:3]
        b = caps[k, 3:6]
        ab = (b - a).astype(np.float32)
        apx, apy, apz = px - a[0], py - a[1], pz - a[2]
        len2 = np.float32(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2])
        if len2 > 0:
            t = (apx * ab[0] + apy * ab[1] + apz * ab[2]) / len2
            t = np.clip(t, np.float32(0), np.float32(1)).astype(np.float32)
        else:
            t = np.zeros(shape, np.float32)
        dx, dy, dz = apx - t * ab[0], apy - t * ab[1], apz - t * ab[2]
        d2 = dx * dx + dy * dy + dz * dz
        hit = (d2 <= np.float32(caps[k, 6] * caps[k, 6])) & ~done
        out[hit] = int(caps[k, 7]) + 1
        done |= hit
    return out
