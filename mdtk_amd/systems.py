"""Synthetic MDTK-style systems and seeded initial velocities.

Positions are returned in Angstrom; ``to_cgs`` converts to the reference's
CGS units (cm, cm/s).  Element ids follow mdtk/consts.hpp:46-56 (H=1, C=12,
Ar=40).  A non-periodic axis carries the reference's NO_PBC sentinel,
1e4 Angstrom (mdtk/consts.hpp:96).

Velocities are Maxwell-Boltzmann at temperature T from a counter-based
Philox-4x32-10 stream (Salmon et al., SC'11), base seed 697860 and one
sub-stream per trajectory, with the mass-weighted mean velocity of the
non-fixed atoms removed (mdtk/AtomsContainer.cxx:205-219,268-278).
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field

import numpy as np

ANG = 1e-8                      # cm, mdtk/consts.cxx:30
EV = 1.6021e-12                 # erg, mdtk/consts.cxx:31
AMU = 1.6605655e-24             # g, mdtk/consts.cxx:28
KB = 8.617343e-5 * EV           # erg/K, mdtk/consts.cxx:32
FS = 1e-15
NO_PBC = 10000.0                # Angstrom
H, C, AR = 1, 12, 40
MASS_AMU = {H: 1.0, C: 12.0, AR: 39.948, 64: 63.546, 108: 107.868, 197: 196.967, 131: 131.293}
BASE_SEED = 697860


@dataclass
class System:
    elem: np.ndarray            # int32 [n]
    pos: np.ndarray             # float64 [n,3], Angstrom
    pbc: np.ndarray             # float64 [3], Angstrom (NO_PBC when open)
    vel: np.ndarray = None      # float64 [n,3], Angstrom/fs
    fixed: np.ndarray = None    # uint8 [n]
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.elem.shape[0])

    def to_cgs(self):
        """(elem, xyz_cm, vel_cm_s, pbc_cm, fixed) in the reference's units."""
        v = self.vel if self.vel is not None else np.zeros_like(self.pos)
        fx = self.fixed if self.fixed is not None else np.zeros(self.n, np.uint8)
        return (np.ascontiguousarray(self.elem, np.int32),
                np.ascontiguousarray(self.pos * ANG),
                np.ascontiguousarray(v * (ANG / FS)),
                np.ascontiguousarray(self.pbc * ANG),
                np.ascontiguousarray(fx, np.uint8))

    def wrapped(self) -> "System":
        p = self.pos.copy()
        for d in range(3):
            if self.pbc[d] != NO_PBC:
                p[:, d] = np.mod(p[:, d], self.pbc[d])
                p[:, d] = np.where(p[:, d] >= self.pbc[d], p[:, d] - self.pbc[d], p[:, d])
        return System(self.elem, p, self.pbc, self.vel, self.fixed, self.name, dict(self.meta))


# --------------------------------------------------------------------- Philox
_M0, _M1 = 0xD2511F53, 0xCD9E8D57
_W0, _W1 = 0x9E3779B9, 0xBB67AE85


def philox4x32(counter: np.ndarray, key: tuple[int, int], rounds: int = 10) -> np.ndarray:
    """Philox-4x32-R on a [n,4] uint32 counter array with a 2x32-bit key."""
    c = counter.astype(np.uint64)
    k0, k1 = np.uint64(key[0] & 0xFFFFFFFF), np.uint64(key[1] & 0xFFFFFFFF)
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(rounds):
        p0 = c[:, 0] * np.uint64(_M0)
        p1 = c[:, 2] * np.uint64(_M1)
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        c = np.stack([(hi1 ^ c[:, 1] ^ k1) & mask, lo1, (hi0 ^ c[:, 3] ^ k0) & mask, lo0], axis=1)
        k0 = (k0 + np.uint64(_W0)) & mask
        k1 = (k1 + np.uint64(_W1)) & mask
    return c.astype(np.uint32)


def normals(n: int, stream: int, seed: int = BASE_SEED) -> np.ndarray:
    """n standard normals from Philox sub-stream `stream` (Box-Muller)."""
    m = (n + 3) // 4
    ctr = np.zeros((m, 4), np.uint32)
    ctr[:, 0] = np.arange(m, dtype=np.uint64) & 0xFFFFFFFF
    ctr[:, 1] = stream
    r = philox4x32(ctr, (seed, 0x6d64746b)).astype(np.float64)
    u = (r + 0.5) / 4294967296.0
    u1, u2 = u[:, 0::2].ravel(), u[:, 1::2].ravel()
    rad = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([rad * np.cos(2 * np.pi * u2), rad * np.sin(2 * np.pi * u2)])
    return z[:n]


def uniforms(n: int, stream: int, seed: int = BASE_SEED) -> np.ndarray:
    m = (n + 3) // 4
    ctr = np.zeros((m, 4), np.uint32)
    ctr[:, 0] = np.arange(m, dtype=np.uint64) & 0xFFFFFFFF
    ctr[:, 1] = stream
    ctr[:, 2] = 0x75
    r = philox4x32(ctr, (seed, 0x6d64746b)).astype(np.float64)
    return ((r.ravel() + 0.5) / 4294967296.0)[:n]


def maxwell_velocities(sys: System, T: float, stream: int = 0, seed: int = BASE_SEED) -> np.ndarray:
    """Angstrom/fs velocities ~ N(0, kT/m) per component, zero net momentum."""
    m = np.array([MASS_AMU[int(e)] for e in sys.elem]) * AMU
    z = normals(3 * sys.n, stream, seed).reshape(sys.n, 3)
    v = z * np.sqrt(KB * T / m)[:, None]          # cm/s
    free = np.ones(sys.n, bool) if sys.fixed is None else (sys.fixed == 0)
    if free.any():
        p = (v[free] * m[free, None]).sum(0) / m[free].sum()
        v[free] -= p
    v[~free] = 0.0
    return v / (ANG / FS)


# ------------------------------------------------------------------ builders
def diamond(nx: int, ny: int, nz: int, a: float = 3.567) -> System:
    """Diamond-cubic carbon, nx*ny*nz conventional cells, 3-D periodic."""
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    pos = ((g[:, None, :] + base[None]) * a).reshape(-1, 3)
    elem = np.full(len(pos), C, np.int32)
    return System(elem, pos, np.array([nx * a, ny * a, nz * a]), name=f"diamond{nx}x{ny}x{nz}")


def polyethylene(nx: int, ny: int, nz: int, a: float = 7.42, b: float = 4.95, c: float = 2.55,
                 chains_per_cell: int = 2, wrap: bool = True) -> System:
    """Orthorhombic polyethylene (Bunn cell), CH2 zig-zag chains along z.

    chains_per_cell=2 is the physical two-chain cell (12 atoms); 1 keeps only the
    corner chain (6 atoms), the basis that gives ~55k atoms at 24x24x16 cells.
    C-C 1.54 A, C-H 1.09 A, C-C-C 112 deg, H-C-H 107 deg (sp3).
    """
    cc, ch = 1.54, 1.09
    half = c / 2.0
    # zig-zag: C at z=0 and z=c/2; lateral offset so that |C-C| = 1.54
    dx = math.sqrt(max(cc * cc - half * half, 0.0))
    chains = [(0.0, 0.0, 0.0)]                      # (x, y, setting angle offset)
    if chains_per_cell == 2:
        chains.append((a / 2, b / 2, 1.0))
    setting = math.radians(42.0)
    atoms_e, atoms_p = [], []
    for (x0, y0, flip) in chains:
        ang = setting if flip == 0.0 else math.pi - setting
        ux, uy = math.cos(ang), math.sin(ang)       # zig-zag plane direction in xy
        for s, zc in ((-1.0, 0.0), (1.0, half)):
            cpos = np.array([x0 + s * ux * dx / 2, y0 + s * uy * dx / 2, zc])
            atoms_e.append(C)
            atoms_p.append(cpos)
            # H pair bisects the outward direction, perpendicular to the chain axis
            out = np.array([s * ux, s * uy, 0.0])
            perp = np.array([-uy, ux, 0.0])
            hch = math.radians(107.0) / 2
            for sg in (-1.0, 1.0):
                d = math.cos(hch) * out + sg * math.sin(hch) * perp
                d = d / np.linalg.norm(d)
                atoms_e.append(H)
                atoms_p.append(cpos + ch * d)
    base_e = np.array(atoms_e, np.int32)
    base_p = np.array(atoms_p)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    shift = g * np.array([a, b, c])
    pos = (shift[:, None, :] + base_p[None]).reshape(-1, 3)
    elem = np.tile(base_e, len(g))
    s = System(elem, pos, np.array([nx * a, ny * b, nz * c]), name=f"pe{nx}x{ny}x{nz}x{chains_per_cell}")
    return s.wrapped() if wrap else s


def graphite(nx: int, ny: int, nz: int, a: float = 2.46, d: float = 3.35) -> System:
    """AB-stacked graphite, orthorhombic 8-atom cell (a, a*sqrt3, 2d), 3-D periodic."""
    b = a * math.sqrt(3.0)
    c = 2 * d
    l0 = np.array([[0, 0], [a / 2, b / 6], [a / 2, b / 2], [0, 2 * b / 3]])
    l1 = l0 + np.array([0, b / 3])
    base = np.concatenate([np.c_[l0, np.zeros(4)], np.c_[l1, np.full(4, d)]])
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    pos = (g[:, None, :] * np.array([a, b, c]) + base[None]).reshape(-1, 3)
    s = System(np.full(len(pos), C, np.int32), pos, np.array([nx * a, ny * b, nz * c]),
               name=f"graphite{nx}x{ny}x{nz}")
    return s.wrapped()


FCC_A = {64: 3.615, 108: 4.086, 197: 4.078}  # Cu, Ag, Au lattice constants, A (Ackland.cxx:304,327,350)


def fcc(elem: int, nx: int, ny: int, nz: int, a: float | None = None, periodic=(True, True, True)) -> System:
    """fcc crystal of one metal, cubic 4-atom cell, periodic along the flagged
    axes (an open axis keeps the block's own extent)."""
    a = FCC_A[elem] if a is None else a
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * a
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    pos = (g[:, None, :] * a + base[None]).reshape(-1, 3) + 0.25 * a
    pbc = np.array([n * a if per else NO_PBC for n, per in zip((nx, ny, nz), periodic)])
    s = System(np.full(len(pos), elem, np.int32), pos, pbc, name=f"fcc{elem}-{nx}x{ny}x{nz}")
    return s.wrapped()


def ethane() -> System:
    """C2H6 (8 atoms, open boundaries): staggered ethane, C-C 1.54 A, C-H 1.09 A."""
    cc, ch = 1.54, 1.09
    th = math.radians(109.47)
    pos = [[0, 0, 0], [0, 0, cc]]
    el = [C, C]
    for k in range(3):
        phi = 2 * math.pi * k / 3
        pos.append([ch * math.sin(th) * math.cos(phi), ch * math.sin(th) * math.sin(phi), ch * math.cos(th)])
        el.append(H)
        phi2 = phi + math.pi / 3
        pos.append([ch * math.sin(th) * math.cos(phi2), ch * math.sin(th) * math.sin(phi2), cc - ch * math.cos(th)])
        el.append(H)
    p = np.array(pos, float) + 20.0
    return System(np.array(el, np.int32), p, np.array([NO_PBC] * 3), name="ethane")


def hydrocarbon_cluster(n_target: int, stream: int, jitter: float = 0.05, seed: int = BASE_SEED) -> System:
    """Random sp3 CH2 chain fragment with ~n_target atoms (C:H ~ 1:2), open boundaries.

    A zig-zag backbone is folded by seeded dihedral choices (trans/gauche);
    each C carries two H (ends carry three), then every coordinate is jittered
    by N(0, jitter) Angstrom.  A dihedral is rejected when the new C comes
    within 2.4 A of a non-neighbour C or 2.0 A of a placed H, or when the H it
    fixes on the previous C come within 2.0 A of a C other than their own or
    1.8 A of another C's H: no H sits inside a foreign atom's REBO cutoff
    (C-H 1.8 A, H-H 1.7 A) before the jitter.  The three dihedrals are tried
    in a random order and the first accepted is kept; when none is, the
    growing end steps back 1, 2, 4, ... C and grows again (at most 4 nC
    times); past that the candidate with the largest clearance is kept.
    """
    cc = 1.54
    nC = max(2, (n_target - 2) // 3)
    ui = _uniform_stream(stream, seed)
    th = math.radians(112.0)
    Cs = [np.zeros(3), np.array([cc, 0, 0]), np.array([cc + cc * math.cos(math.pi - th), cc * math.sin(math.pi - th), 0])]
    Hs = [_hydrogens(Cs, 0, 3), _hydrogens(Cs, 1, 3)]  # H of every C whose neighbours are placed
    carr = np.zeros((max(nC, 3), 3))
    carr[:3] = Cs
    harr = np.zeros((3 * max(nC, 3) + 2, 3))
    nh = 5
    harr[:nh] = np.concatenate(Hs)
    dihs = (math.pi, math.pi / 3, -math.pi / 3)
    orders = list(itertools.permutations(range(3)))
    backtracks, fails, longest = 0, 0, 3
    while len(Cs) < nC:
        k = len(Cs)
        best, best_clear = None, -np.inf
        for d in orders[int(next(ui) * 6) % 6]:  # the dihedrals in a random order
            p = _place(Cs[-3], Cs[-2], Cs[-1], cc, th, dihs[d])
            h_prev = _hydrogens((Cs[-2], Cs[-1], p), 1, 3)
            clear = min(np.sqrt(((carr[:k - 2] - p) ** 2).sum(1).min()) - 2.4,
                        np.sqrt(((harr[:nh] - p) ** 2).sum(1).min()) - 2.0,
                        np.sqrt(((carr[None, :k - 1] - h_prev[:, None]) ** 2).sum(2).min()) - 2.0,
                        np.sqrt(((harr[None, :nh] - h_prev[:, None]) ** 2).sum(2).min()) - 1.8)
            if clear > best_clear:
                best, best_clear = (p, h_prev), clear
            if clear > 0:
                break
        if best_clear <= 0 and k > 3 and backtracks < 4 * nC:
            # a trapped end: take back the last 1, 2, 4, ... C (with the H
            # they fixed) and grow again
            backtracks += 1
            fails += 1
            for _ in range(min(1 << min(fails - 1, 10), k - 3)):
                Cs.pop()
                Hs.pop()
                nh -= 2
            continue
        Cs.append(best[0])
        Hs.append(best[1])
        carr[k] = best[0]
        harr[nh:nh + 2] = best[1]
        nh += 2
        if len(Cs) > longest:
            longest, fails = len(Cs), 0
    Cs = Cs[:nC]
    Hs = Hs[:nC - 1] + [_hydrogens(Cs, nC - 1, nC)]
    el, pos = [], []
    for cpos, hs in zip(Cs, Hs):
        el.append(C)
        pos.append(cpos)
        el += [H] * len(hs)
        pos += list(hs)
    pos = np.array(pos)
    pos += jitter * normals(pos.size, stream + 1000003, seed).reshape(pos.shape)
    pos -= pos.min(0)
    pos += 10.0
    return System(np.array(el, np.int32), pos, np.array([NO_PBC] * 3), name=f"cluster{stream}")


def _uniform_stream(stream, seed):
    """Uniforms of one Philox stream, 4096 at a time, for as long as asked."""
    block = 0
    while True:
        yield from uniforms(4096, stream + 7919 * block, seed)
        block += 1


def _hydrogens(Cs, i, nC):
    """The H of backbone C i (of the first nC): two in the H-C-H plane
    normal to the C-C-C plane, H-C-H 109.47 deg; three, staggered, at a chain
    end.  C-H 1.09 A."""
    ch, hch = 1.09, math.radians(109.47)
    cpos = Cs[i]
    nb = [Cs[j] for j in (i - 1, i + 1) if 0 <= j < nC]
    bis = -sum((q - cpos) / np.linalg.norm(q - cpos) for q in nb)
    if np.linalg.norm(bis) < 1e-6:
        bis = np.array([0, 0, 1.0])
    bis /= np.linalg.norm(bis)
    ax = (nb[0] - cpos) if len(nb) == 1 else (nb[1] - nb[0])
    perp = np.cross(bis, ax)
    for trial in ([1.0, 0, 0], [0, 1.0, 0]):
        if np.linalg.norm(perp) > 1e-6:
            break
        perp = np.cross(bis, trial)
    perp /= np.linalg.norm(perp)
    if len(nb) == 2:
        dirs = [math.cos(hch / 2) * bis + s * math.sin(hch / 2) * perp for s in (-1, 1)]
    else:
        axn = -(nb[0] - cpos) / np.linalg.norm(nb[0] - cpos)
        e1 = perp
        e2 = np.cross(axn, e1)
        dirs = [math.cos(math.radians(70.5)) * axn + math.sin(math.radians(70.5)) *
                (math.cos(2 * math.pi * k / 3) * e1 + math.sin(2 * math.pi * k / 3) * e2) for k in range(3)]
    return np.array([cpos + ch * d / np.linalg.norm(d) for d in dirs])


def _place(a, b, c, bond, theta, dih):
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    n /= np.linalg.norm(n)
    m = np.cross(n, bc)
    d2 = np.array([-bond * math.cos(theta), bond * math.sin(theta) * math.cos(dih), bond * math.sin(theta) * math.sin(dih)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def pe_slab(layers: int, lateral: tuple[int, int] = (8, 3)) -> System:
    """Polyethylene slab with the chains in the surface plane: chains run along
    x (periodic, lateral[0] chain periods c), the a axis of the Bunn cell lies
    along y (periodic, lateral[1] cells), and `layers` cells of b stack along
    z, which is open (NO_PBC) with a free top and bottom surface.
    12*lateral[0]*lateral[1] atoms per layer."""
    nc, na = lateral
    pe = polyethylene(na, layers, nc, wrap=False)   # chains along the pe's z
    a, b, c = pe.pbc
    raw = pe.pos
    pos = np.stack([raw[:, 2], raw[:, 0], raw[:, 1]], 1)  # x <- chain axis, y <- a, z <- b
    pos[:, 2] -= pos[:, 2].min()
    s = System(pe.elem.copy(), pos, np.array([c, a, NO_PBC]), name=f"peslab{nc}x{na}x{layers}")
    return s.wrapped()


def bombard_target(n_target: int, stream: int, energy_ev: float = 100.0, T: float = 300.0,
                   lateral: tuple[int, int] = (8, 3), height: float = 8.0, seed: int = BASE_SEED) -> System:
    """One bombardment trajectory: a pe_slab with about n_target atoms (whole
    layers), the bottom chain row fixed (Atom::fix, mdtk/Atom.hpp:58), 300 K
    Maxwell-Boltzmann velocities, and one Ar ion `height` Angstrom above the
    top surface at a seeded random lateral point, moving along -z with
    `energy_ev` (normal incidence)."""
    per_layer = 12 * lateral[0] * lateral[1]
    layers = max(1, int(round(n_target / per_layer)))
    s = pe_slab(layers, lateral)
    z = s.pos[:, 2]
    # bottom chain row: the corner chain (first 6 atoms of a cell) of layer 0
    idx = np.arange(s.n)
    cell = idx // 12
    layer = (cell // lateral[0]) % layers
    fixed = ((layer == 0) & (idx % 12 < 6)).astype(np.uint8)
    s.fixed = fixed
    s.vel = maxwell_velocities(s, T, stream=stream, seed=seed)
    u = uniforms(2, stream + 2000003, seed)
    ion = np.array([[u[0] * s.pbc[0], u[1] * s.pbc[1], z.max() + height]])
    v_ion = math.sqrt(2.0 * energy_ev * EV / (MASS_AMU[AR] * AMU)) / (ANG / FS)
    elem = np.concatenate([s.elem, [AR]]).astype(np.int32)
    pos = np.concatenate([s.pos, ion])
    vel = np.concatenate([s.vel, [[0.0, 0.0, -v_ion]]])
    fx = np.concatenate([fixed, [0]]).astype(np.uint8)
    return System(elem, pos, s.pbc.copy(), vel, fx, name=f"bombard{stream}",
                  meta=dict(layers=layers, ion_energy_ev=energy_ev))


def bombard_batch(ntraj: int = 256, nmin: int = 2000, nmax: int = 12000, **kw) -> list[System]:
    """ntraj bombardment trajectories with slab sizes ~U[nmin, nmax] (seeded,
    rounded to whole layers); trajectory t uses Philox sub-stream t."""
    u = uniforms(ntraj, 4000037)
    return [bombard_target(int(nmin + u[t] * (nmax - nmin)), t, **kw) for t in range(ntraj)]


def jitter(sys: System, sigma: float, stream: int) -> System:
    p = sys.pos + sigma * normals(sys.pos.size, stream).reshape(sys.pos.shape)
    s = System(sys.elem, p, sys.pbc, sys.vel, sys.fixed, sys.name, dict(sys.meta))
    return s.wrapped()
