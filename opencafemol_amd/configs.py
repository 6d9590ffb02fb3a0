"""Seeded synthetic workloads, built in OpenMM units.

* ``go_chain``      -- Calpha Go chain: harmonic bonds and angles, Gaussian
                       dihedrals, 12-10 native contacts, excluded volume and
                       Debye-Hueckel on a charged subset, open boundaries.
* ``hps_slab``      -- HPS (Ashbaugh-Hatch) chains in a periodic box, packed
                       in a central slab: harmonic bonds, LJ repulsive +
                       lambda-weighted LJ attractive (switched), Debye-Hueckel.
* ``go_replicas``   -- many Go chains of uneven length on a temperature ladder.

Parameters are given the way the reference's TOML readers receive them
(Angstrom, kcal/mol, the E = k (x - x0)^2 convention for harmonic terms) and
converted here exactly as those readers convert them
(src/input/ReadTOMLForceFieldGenerator.cpp:57-64 bonds, :397-398 angles,
:195-198 contacts, :531-536 dihedrals, :1066-1121 excluded volume,
:2112-2237 Lennard-Jones).
"""
import math

import numpy as np

from . import abi
from .model import ForceSpec, IntegratorSpec, SystemSpec

KJ_PER_KCAL = 4.184
NM_PER_A = 0.1
NA = 6.02214076e23
KB = 1.380649e-23
EPS0 = 8.854187817e-12 * 1.0e3 * 1.0e-9 / NA   # C^2 mol / kJ / nm
QE = 1.6021766208e-19
CAFETIME = math.sqrt(1.0 / KJ_PER_KCAL) * 0.1   # ps (src/util/Constants.hpp:23)


def dielectric_water(T, C):
    """DebyeHuckelForceFieldGenerator.hpp:36-41."""
    return (249.4 - 0.788 * T + 7.2e-4 * T * T) * (1.0 - 2.551e-1 * C + 5.151e-2 * C * C - 6.889e-3 * C ** 3)


def debye_huckel_constants(ionic_strength, temperature, cutoff_ratio):
    """(K = 1/(4 pi eps0 epsk), Debye length, cutoff correction, absolute cutoff)
    as computed by DebyeHuckelForceFieldGenerator.cpp:10-51."""
    if not ionic_strength > 0:
        raise ValueError("ionic strength must be larger than zero")
    epsk = dielectric_water(temperature, ionic_strength)
    eps0 = EPS0 / QE / QE
    K = 1.0 / (4.0 * math.pi * eps0 * epsk)
    I = ionic_strength * 1.0e-24
    lam = math.sqrt((eps0 * epsk * KB * 1.0e-3 * temperature) / (2.0 * I))
    rc = lam * cutoff_ratio
    corr = math.exp(-cutoff_ratio) / rc
    return K, lam, corr, rc


def ionic_strength_for_length(lam, temperature):
    """Ionic strength [M] whose Debye length at `temperature` is `lam` nm."""
    lo, hi = 1e-6, 5.0
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        if debye_huckel_constants(mid, temperature, 1.0)[1] > lam:
            lo = mid
        else:
            hi = mid
    return math.sqrt(lo * hi)


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def saw_chain(n, rng, bond=0.38, min_dist=0.40, start=(0.0, 0.0, 0.0), angle_range=(75.0, 140.0)):
    """Self-avoiding random walk of n beads (nm): every new bead is `bond` from
    the previous one, at least `min_dist` from all earlier beads, and makes a
    bond angle within `angle_range` (degrees) at the previous bead, the span of
    C-alpha pseudo-bond angles.  Native angles far from 180 degrees keep the
    chain away from collinear triples, where dihedral gradients diverge."""
    cos_hi, cos_lo = np.cos(np.radians(angle_range[0])), np.cos(np.radians(angle_range[1]))
    x = np.zeros((n, 3))
    x[0] = start
    i = 1
    while i < n:
        for _ in range(200):
            cand = x[i - 1] + bond * _unit(rng, 1)[0]
            if i >= 2:
                u, w = x[i - 2] - x[i - 1], cand - x[i - 1]
                ca = np.dot(u, w) / (np.linalg.norm(u) * np.linalg.norm(w))
                if not (cos_lo <= ca <= cos_hi):
                    continue
            if i < 2 or np.min(np.linalg.norm(x[: i - 1] - cand, axis=1)) >= min_dist:
                x[i] = cand
                i += 1
                break
        else:  # stuck: restart from a shorter prefix
            i = max(1, i - 10)
    return x


def angles_of(x, tri):
    a = x[tri[:, 0]] - x[tri[:, 1]]
    c = x[tri[:, 2]] - x[tri[:, 1]]
    return np.arctan2(np.linalg.norm(np.cross(a, c), axis=1), np.einsum("ij,ij->i", a, c))


def dihedrals_of(x, quad):
    b1 = x[quad[:, 1]] - x[quad[:, 0]]
    b2 = x[quad[:, 2]] - x[quad[:, 1]]
    b3 = x[quad[:, 3]] - x[quad[:, 2]]
    n1 = np.cross(b1, b2)
    n2 = np.cross(b2, b3)
    return np.arctan2(np.linalg.norm(b2, axis=1) * np.einsum("ij,ij->i", b1, n2), np.einsum("ij,ij->i", n1, n2))


def chain_exclusions(chains, within):
    """Pairs (i<j) within `within` bonds along linear chains, the set
    Topology::ignore_list_within_edge (src/Topology.cpp:75-110) yields for
    linear bond graphs.  `chains` is a list of (start, length)."""
    out = []
    for s, n in chains:
        for d in range(1, within + 1):
            i = np.arange(s, s + n - d)
            out.append(np.stack([i, i + d], axis=1))
    return np.concatenate(out).astype(np.int32) if out else np.zeros((0, 2), np.int32)


def go_chain(n=62, seed=0, temperature=300.0, friction=0.5, dt=0.01, charged_fraction=0.2,
             mass=110.0, angle_k=20.0, contact_eps=0.3, exv_eps=0.6, exv_radius=2.0,
             dh_length=1.0, dh_cutoff_ratio=3.0, exv_ratio=2.0):
    """A Go-model Calpha chain grown as a seeded self-avoiding walk.

    Native angles are held in [75, 140] degrees (saw_chain) and the angle
    stiffness is the Clementi-type Go value, 20 kcal/mol/rad^2.  An
    unconstrained walk has native angles near 180 degrees; replicas near 400 K
    then reach collinear triples within ~10,000 steps, where the Gaussian
    dihedral's gradient (which scales as 1/sin(angle)) ejects a bead and the
    chain explodes."""
    rng = np.random.default_rng(seed)
    x = saw_chain(n, rng)
    sysm = SystemSpec(masses=np.full(n, mass))
    bonds = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1).astype(np.int32)
    r0 = np.linalg.norm(x[bonds[:, 1]] - x[bonds[:, 0]], axis=1)
    kb = 110.0 * KJ_PER_KCAL / (NM_PER_A ** 2) * 2.0
    sysm.add(ForceSpec(abi.HARMONIC_BOND, "HarmonicBond", indices=bonds,
                       params=np.stack([r0, np.full(n - 1, kb)], axis=1)))
    tri = np.stack([np.arange(n - 2), np.arange(1, n - 1), np.arange(2, n)], axis=1).astype(np.int32)
    th0 = angles_of(x, tri)
    sysm.add(ForceSpec(abi.HARMONIC_ANGLE, "HarmonicAngle", indices=tri,
                       params=np.stack([th0, np.full(n - 2, angle_k * KJ_PER_KCAL * 2.0)], axis=1)))
    quad = np.stack([np.arange(n - 3), np.arange(1, n - 2), np.arange(2, n - 1), np.arange(3, n)], axis=1).astype(np.int32)
    ph0 = dihedrals_of(x, quad)
    sysm.add(ForceSpec(abi.GAUSSIAN_DIHEDRAL, "GaussianDihedral", indices=quad,
                       params=np.stack([np.full(n - 3, -1.0 * KJ_PER_KCAL), ph0, np.full(n - 3, 0.15)], axis=1)))
    ii, jj = np.triu_indices(n, k=4)
    d = np.linalg.norm(x[ii] - x[jj], axis=1)
    sel = d < 0.65
    contacts = np.stack([ii[sel], jj[sel]], axis=1).astype(np.int32)
    if len(contacts):
        sysm.add(ForceSpec(abi.GO_CONTACT, "GoContact", indices=contacts,
                           params=np.stack([np.full(len(contacts), contact_eps * KJ_PER_KCAL), d[sel]], axis=1)))
    excl_exv = np.concatenate([chain_exclusions([(0, n)], 3), contacts]) if len(contacts) else chain_exclusions([(0, n)], 3)
    radius = exv_radius * NM_PER_A
    sysm.add(ForceSpec(abi.EXCLUDED_VOLUME, "ExcludedVolume", particle_params=np.full((n, 1), radius),
                       globals=np.array([exv_eps * KJ_PER_KCAL]), cutoff=exv_ratio * (radius + radius),
                       exclusions=excl_exv))
    nq = int(round(charged_fraction * n))
    if nq > 0:
        charged = rng.choice(n, size=nq, replace=False)
        q = np.zeros(n)
        q[charged] = rng.choice([-1.0, 1.0], size=nq)
        cat = np.full(n, -1, np.int32)
        cat[charged] = 0
        I = ionic_strength_for_length(dh_length, temperature)
        K, lam, corr, rc = debye_huckel_constants(I, temperature, dh_cutoff_ratio)
        sysm.add(ForceSpec(abi.DEBYE_HUCKEL, "DebyeHuckel", particle_params=q.reshape(n, 1),
                           globals=np.array([K, lam, corr]), category=cat, n_categories=1, cutoff=rc,
                           exclusions=chain_exclusions([(0, n)], 3)))
    integ = IntegratorSpec(temperature, friction, dt, seed=(seed * 2654435761 + 12345) & 0xFFFFFFFFFFFF)
    return sysm, integ, x


def lattice_chains(n_chains, chain_len, box, slab_z, rng, spacing=0.5):
    """Chains as self-avoiding walks on a cubic lattice (spacing nm) filling the
    full x-y extent of `box` and a central slab of thickness `slab_z`; x-y wrap
    periodically.  Returns positions (n_chains*chain_len, 3)."""
    nx = int(box[0] // spacing)
    ny = int(box[1] // spacing)
    nz = int(slab_z // spacing)
    if n_chains * chain_len > 0.6 * nx * ny * nz:
        raise ValueError("slab too thin for the requested beads")
    occ = np.zeros((nx, ny, nz), dtype=bool)
    dirs = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    out = np.zeros((n_chains, chain_len, 3), dtype=np.int64)
    active = np.arange(n_chains)

    def free_sites(k):
        cand = np.stack([rng.integers(0, nx, 4 * k), rng.integers(0, ny, 4 * k), rng.integers(0, nz, 4 * k)], 1)
        return cand[~occ[cand[:, 0], cand[:, 1], cand[:, 2]]]

    starts = []
    while len(starts) < n_chains:
        for c in free_sites(n_chains):
            t = tuple(c)
            if not occ[t]:
                occ[t] = True
                starts.append(c)
                if len(starts) == n_chains:
                    break
    out[:, 0] = np.array(starts)
    length = np.ones(n_chains, dtype=np.int64)
    stalls = np.zeros(n_chains, dtype=np.int64)
    while True:
        grow = np.nonzero(length < chain_len)[0]
        if len(grow) == 0:
            break
        for c in grow:
            cur = out[c, length[c] - 1]
            order = rng.permutation(6)
            placed = False
            for k in order:
                nxt = cur + dirs[k]
                nxt[0] %= nx
                nxt[1] %= ny
                if nxt[2] < 0 or nxt[2] >= nz or occ[nxt[0], nxt[1], nxt[2]]:
                    continue
                occ[nxt[0], nxt[1], nxt[2]] = True
                out[c, length[c]] = nxt
                length[c] += 1
                placed = True
                break
            if placed:
                continue
            stalls[c] += 1
            if stalls[c] > 100:
                # trapped in a pocket shorter than the chain (the 14,000-bead
                # slab sees at most 6 dead ends per chain): restart it from a free site
                for b in range(length[c]):
                    p = out[c, b]
                    occ[p[0], p[1], p[2]] = False
                cand = free_sites(1)
                while len(cand) == 0:
                    cand = free_sites(1)
                start = cand[0]
                occ[start[0], start[1], start[2]] = True
                out[c, 0] = start
                length[c] = 1
                stalls[c] = 0
                continue
            # dead end: release the last 5 beads and regrow
            back = min(5, length[c] - 1)
            for b in range(back):
                p = out[c, length[c] - 1]
                occ[p[0], p[1], p[2]] = False
                length[c] -= 1
    del active
    z0 = 0.5 * (box[2] - nz * spacing)
    pos = out.reshape(-1, 3).astype(np.float64) * spacing + np.array([0.25 * spacing, 0.25 * spacing, z0])
    return pos


# 20 residue types of the HPS slab: sigma ~ U[0.45, 0.80] nm, lambda ~ U[0, 1]
def hps_types(rng, n_types=20):
    return rng.uniform(0.45, 0.80, n_types), rng.uniform(0.0, 1.0, n_types)


def hps_slab(n_chains=100, chain_len=140, box=(15.0, 15.0, 150.0), slab_z=20.0, seed=0, temperature=300.0,
             friction=0.5, dt=0.01, mass=110.0, eps=0.2, cutoff=2.0, dh_length=1.0, dh_cutoff=3.5,
             charged_fraction=(0.15, 0.15), bond_k=110.4, positions=None):
    """HPS slab: chains packed on a lattice in a central slab of a periodic box
    (or at `positions`, when given)."""
    rng = np.random.default_rng(seed)
    n = n_chains * chain_len
    x = lattice_chains(n_chains, chain_len, box, slab_z, rng) if positions is None else np.asarray(positions)
    sig_t, lam_t = hps_types(rng)
    types = rng.integers(0, len(sig_t), n)
    sysm = SystemSpec(masses=np.full(n, mass), box=tuple(box))
    chains = [(c * chain_len, chain_len) for c in range(n_chains)]
    bonds = np.concatenate([np.stack([np.arange(s, s + L - 1), np.arange(s + 1, s + L)], 1) for s, L in chains]).astype(np.int32)
    kb = bond_k * KJ_PER_KCAL / (NM_PER_A ** 2) * 2.0
    sysm.add(ForceSpec(abi.HARMONIC_BOND, "HarmonicBond", use_periodic=True, indices=bonds,
                       params=np.stack([np.full(len(bonds), 0.38), np.full(len(bonds), kb)], 1)))
    excl1 = chain_exclusions(chains, 1)
    e_rep = np.full(n, eps * KJ_PER_KCAL)
    e_att = eps * KJ_PER_KCAL * lam_t[types]
    sig = sig_t[types]
    sysm.add(ForceSpec(abi.LJ_REPULSIVE, "LennardJonesRepulsive", use_periodic=True,
                       particle_params=np.stack([e_rep, sig], 1), cutoff=cutoff, exclusions=excl1))
    sysm.add(ForceSpec(abi.LJ_ATTRACTIVE, "LennardJonesAttractive", use_periodic=True,
                       particle_params=np.stack([e_att, sig], 1), cutoff=cutoff, switch_distance=0.9 * cutoff,
                       exclusions=excl1))
    npos, nneg = int(round(charged_fraction[0] * n)), int(round(charged_fraction[1] * n))
    idx = rng.permutation(n)
    q = np.zeros(n)
    q[idx[:npos]] = 1.0
    q[idx[npos:npos + nneg]] = -1.0
    cat = np.where(q != 0, 0, -1).astype(np.int32)
    ratio = dh_cutoff / dh_length
    I = ionic_strength_for_length(dh_length, temperature)
    K, lam, corr, rc = debye_huckel_constants(I, temperature, ratio)
    sysm.add(ForceSpec(abi.DEBYE_HUCKEL, "DebyeHuckel", use_periodic=True, particle_params=q.reshape(n, 1),
                       globals=np.array([K, lam, corr]), category=cat, n_categories=1, cutoff=rc,
                       exclusions=chain_exclusions(chains, 3)))
    integ = IntegratorSpec(temperature, friction, dt, seed=(seed * 2654435761 + 777) & 0xFFFFFFFFFFFF)
    return sysm, integ, x


def temperature_ladder(n, t_lo=280.0, t_hi=400.0):
    return [t_lo * (t_hi / t_lo) ** (r / max(n - 1, 1)) for r in range(n)]


def go_replicas(n_replicas=256, n_min=50, n_max=500, seed=0):
    """Replica ensemble: Go chains with bead counts ~ U[n_min, n_max] on a
    geometric 280-400 K ladder, each with its own seed."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, n_replicas)
    temps = temperature_ladder(n_replicas)
    out = []
    for r in range(n_replicas):
        out.append(go_chain(int(sizes[r]), seed=seed * 100003 + r, temperature=temps[r]))
    return out


def hps_slab_of_size(n_beads, seed=0, chain_len=140):
    """HPS slab of about n_beads beads (chains of chain_len) at the density of
    the 14,000-bead slab: x-y extent scaled with the bead count (at least 8 nm,
    so the 3.5 nm Debye-Hueckel cutoff stays under half the box), z 150 nm."""
    n_chains = max(1, int(round(n_beads / chain_len)))
    xy = max(8.0, 15.0 * float(np.sqrt(n_chains * chain_len / 14000.0)))
    return hps_slab(n_chains=n_chains, chain_len=chain_len, box=(xy, xy, 150.0), slab_z=20.0, seed=seed)


def serpentine_chains(n_chains, chain_len, box, spacing=(0.38, 0.95, 0.95)):
    """Chains laid along one boustrophedon path through the sites of an
    orthorhombic lattice (spacing nm along x, y, z) that fills the x-y extent of
    `box` and as many z layers as the beads need, centred in z.  Steps along a
    row are spacing[0] (the bond length, so chains start unstretched); row and
    layer turns are one spacing[1] or spacing[2] step.  One site is left empty
    between consecutive chains, so unbonded beads start at least
    2 * spacing[0] apart.  No two beads share a site.  Vectorised: O(N) for a
    million beads.  Returns (N, 3) nm."""
    sx, sy, sz = spacing
    n = n_chains * chain_len
    nx, ny = int(box[0] // sx), int(box[1] // sy)
    per_layer = nx * ny
    n_sites = n + n_chains - 1
    nz = -(-n_sites // per_layer)
    if nz * sz > box[2]:
        raise ValueError("box too small for the requested beads")
    k = np.arange(n)
    k = k + k // chain_len
    layer = k // per_layer
    r = k % per_layer
    row = r // nx
    col = r % nx
    col = np.where((layer * ny + row) % 2 == 1, nx - 1 - col, col)   # snake along rows
    row = np.where(layer % 2 == 1, ny - 1 - row, row)                # and back through the next layer
    z0 = 0.5 * (box[2] - nz * sz)
    return np.stack([(col + 0.5) * sx, (row + 0.5) * sy, z0 + (layer + 0.5) * sz], 1).astype(np.float64)


def hps_condensate(n_chains=1000, chain_len=1000, box=(60.0, 60.0, 600.0), seed=0, spacing=(0.38, 0.95, 0.95)):
    """Large HPS condensate: the slab's parameter distributions (20 types,
    sigma ~ U[0.45, 0.80] nm, lambda ~ U[0, 1], 15%/15% charges, AH cutoff
    2.0 nm, DH 1 nm / 3.5 nm) on serpentine-packed chains in a periodic box,
    about 2.9 beads/nm^3 in the dense layer (the 14,000-bead slab's lattice
    start holds 3.1)."""
    x = serpentine_chains(n_chains, chain_len, box, spacing)
    s, it, _ = hps_slab(n_chains=n_chains, chain_len=chain_len, box=box, seed=seed, positions=x)
    return s, it, x


def hps_open_cluster(n_chains=150, chain_len=140, extent=(12.0, 12.0), seed=0):
    """HPS chains (the slab's parameter distributions) with open boundaries:
    serpentine-packed in a column of the given x-y extent (nm), with every
    force non-periodic and no box -- an open-boundary system large enough
    to need a cell grid (21,000 beads by default)."""
    box = (extent[0], extent[1], 1.0e4)
    x = serpentine_chains(n_chains, chain_len, box)
    x[:, 2] -= x[:, 2].min()
    s, it, _ = hps_slab(n_chains=n_chains, chain_len=chain_len, box=box, seed=seed, positions=x)
    s.box = None
    for f in s.forces:
        f.use_periodic = False
    return s, it, x


def served_job_mix(n_jobs=24, seed=0, steps=(10000, 50000), frame_interval=10000, go_fraction=0.6):
    """Job stream: Go chains of 50-500 beads and HPS slabs of 2,000-20,000
    beads, each running a whole number of frame intervals drawn uniformly
    from [steps[0], steps[1]].  round(go_fraction * n_jobs) of the jobs are Go
    chains, in a random order, so a short stream holds the stated mix."""
    rng = np.random.default_rng(seed)
    fi = frame_interval
    n_go = int(round(go_fraction * n_jobs))
    is_go = rng.permutation(np.arange(n_jobs) < n_go)
    jobs = []
    for k in range(n_jobs):
        if is_go[k]:
            s, it, x = go_chain(int(rng.integers(50, 501)), seed=seed * 1000 + k)
            kind = "go"
        else:
            s, it, x = hps_slab_of_size(int(rng.integers(2000, 20001)), seed=seed * 1000 + k)
            kind = "hps"
        nsteps = int(rng.integers(steps[0] // fi, steps[1] // fi + 1)) * fi
        jobs.append({"kind": kind, "system": s, "integrator": it, "positions": x, "nsteps": nsteps,
                     "frame_interval": fi})
    return jobs


def aicg2_chain(n=56, seed=0, temperature=300.0, periodic_box=None):
    """Go chain plus the AICG2+ local terms the sh3 sample uses: Gaussian 1-3
    bonds, flexible local angles and dihedrals on random residue labels."""
    from . import aicg2
    from .aicg2_tables import FLA_Y
    sysm, integ, x = go_chain(n, seed=seed, temperature=temperature)
    rng = np.random.default_rng(seed + 99)
    residues = list(rng.choice(sorted(FLA_Y), n))
    pairs = np.stack([np.arange(n - 2), np.arange(2, n)], axis=1).astype(np.int32)
    v0 = np.linalg.norm(x[pairs[:, 1]] - x[pairs[:, 0]], axis=1)
    kk = -rng.uniform(1.0, 5.0, n - 2) * KJ_PER_KCAL
    sysm.add(ForceSpec(abi.GAUSSIAN_BOND, "GaussianBond", indices=pairs,
                       params=np.stack([kk, v0, np.full(n - 2, 0.015)], axis=1)))
    tri = np.stack([np.arange(n - 2), np.arange(1, n - 1), np.arange(2, n)], axis=1)
    for f in aicg2.flexible_local_angle_forces(tri, residues, 1.0):
        sysm.add(f)
    quad = np.stack([np.arange(n - 3), np.arange(1, n - 2), np.arange(2, n - 1), np.arange(3, n)], axis=1)
    for f in aicg2.flexible_local_dihedral_forces(quad, residues, 1.0):
        sysm.add(f)
    return sysm, integ, x


def wca_cutoff(sigmas):
    """(largest + second largest sigma) / 2 * 2^(1/6), both tracked with <=
    (WeeksChandlerAndersenForceFieldGenerator.cpp:49-66,101-103)."""
    mx = second = np.finfo(float).tiny
    for s in sigmas:
        if mx <= s:
            second, mx = mx, s
        elif second <= s:
            second = s
    return (mx + second) * 0.5 * 2.0 ** (1.0 / 6.0)


def lipid_patch(n_lipids=216, box=(9.0, 9.0, 9.0), seed=0, temperature=300.0, friction=0.5, dt=0.01,
                mass=150.0, spacing=1.5):
    """Three-bead lipids (head + two tails) in a periodic box, with the
    membrane potentials of the iSoLF family: WCA between every bead (per-bead
    sigma and epsilon), iSoLF attraction between tail beads, and two uniform
    (table) potentials between named bead types: WCA between heads and tails
    (H-T) and a switched LJ attraction among tails (T-T).  Sizes follow the
    reference's POPC comments (ReadTOMLForceFieldGenerator.cpp:1423-1432):
    sigma_H 4.62 A, sigma_T 7.11 A, epsilon 0.416 kcal/mol, omega 9.87 A.
    Lipid centres sit on a jittered cubic lattice, each lipid straight along
    a random direction, 0.55 nm between bonded beads."""
    rng = np.random.default_rng(seed)
    per = int(round(n_lipids ** (1.0 / 3.0)))
    grid = np.stack(np.meshgrid(*[np.arange(per)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n_lipids]
    centers = (grid + 0.5) * spacing + rng.uniform(-0.1, 0.1, size=grid.shape)
    dirs = _unit(rng, n_lipids)
    x = np.concatenate([centers[:, None, :] + dirs[:, None, :] * (np.array([-0.55, 0.0, 0.55])[None, :, None])],
                       axis=0).reshape(-1, 3) % np.asarray(box)
    n = 3 * n_lipids
    is_head = np.tile([True, False, False], n_lipids)
    sysm = SystemSpec(masses=np.full(n, mass), box=tuple(box))
    bonds = np.concatenate([np.stack([3 * np.arange(n_lipids) + k, 3 * np.arange(n_lipids) + k + 1], 1)
                            for k in (0, 1)]).astype(np.int32)
    sysm.add(ForceSpec(abi.HARMONIC_BOND, "HarmonicBond", use_periodic=True, indices=bonds,
                       params=np.tile([0.55, 2.0 * 20.0 * KJ_PER_KCAL * 100.0], (len(bonds), 1))))
    ang = np.stack([3 * np.arange(n_lipids), 3 * np.arange(n_lipids) + 1, 3 * np.arange(n_lipids) + 2], 1).astype(np.int32)
    sysm.add(ForceSpec(abi.HARMONIC_ANGLE, "HarmonicAngle", use_periodic=True, indices=ang,
                       params=np.tile([np.pi, 2.0 * 2.0 * KJ_PER_KCAL], (n_lipids, 1))))
    excl = chain_exclusions([(3 * k, 3) for k in range(n_lipids)], 1)
    sig = np.where(is_head, 4.62215, 7.111) * NM_PER_A
    eps = np.where(is_head, 0.30, 0.416) * KJ_PER_KCAL
    sysm.add(ForceSpec(abi.WCA, "WeeksChandlerAndersen", use_periodic=True,
                       particle_params=np.stack([sig, eps], 1), cutoff=wca_cutoff(sig), exclusions=excl))
    # iSoLF attraction among tails only (heads carry no parameters)
    om = np.full(n, 9.867 * NM_PER_A)
    cat = np.where(is_head, -1, 0).astype(np.int32)
    tails = ~is_head
    isolf_rc = wca_cutoff(sig[tails]) + om[0]
    sysm.add(ForceSpec(abi.ISOLF_ATTRACTIVE, "iSoLFAttractive", use_periodic=True,
                       particle_params=np.stack([sig, eps, om], 1), category=cat, n_categories=1,
                       cutoff=isolf_rc, exclusions=excl))
    # table potentials: WCA between heads (H) and tails (T), LJ attraction T-T
    ht = np.where(is_head, 0, 1).astype(np.int32)
    s_ht = 5.5 * NM_PER_A
    sysm.add(ForceSpec(abi.UNIFORM_WCA, "UniformWeeksChandlerAndersen", use_periodic=True,
                       globals=np.array([s_ht, 0.3 * KJ_PER_KCAL]), category=ht, n_categories=2,
                       allow=np.array([[0, 1], [1, 0]], np.uint8), cutoff=s_ht * 2.0 ** (1.0 / 6.0),
                       exclusions=excl))
    s_tt, ratio = 6.0 * NM_PER_A, 2.5
    tt = np.where(is_head, -1, 0).astype(np.int32)
    sysm.add(ForceSpec(abi.UNIFORM_LJ_ATTRACTIVE, "UniformLennardJonesAttractive", use_periodic=True,
                       globals=np.array([s_tt, 0.2 * KJ_PER_KCAL]), category=tt, n_categories=1,
                       cutoff=s_tt * ratio, switch_distance=0.9 * s_tt * ratio, exclusions=excl))
    return sysm, IntegratorSpec(temperature, friction, dt, seed), x


def capped_go_chain(n=60, seed=0, capping_ratio=0.75, temperature=300.0):
    """A Go chain (go_chain) whose native contacts are capped 12-10 contacts
    and whose dihedrals carry cosine terms besides the Gaussian ones, in the
    TOML reader's conventions: k in kcal/mol, v0 + pi as t0 and integer n
    (ReadTOMLForceFieldGenerator.cpp:226-277, :562-622)."""
    sysm, integ, x = go_chain(n, seed=seed, temperature=temperature)
    rng = np.random.default_rng(seed + 7)
    for f in sysm.forces:
        if f.kind == abi.GO_CONTACT:
            f.kind = abi.CAPPED_GO_CONTACT
            f.name = "CappedGoContact"
            f.params = np.concatenate([f.params, np.full((f.n_terms, 1), capping_ratio)], axis=1)
    quad = np.stack([np.arange(n - 3), np.arange(1, n - 2), np.arange(2, n - 1), np.arange(3, n)], 1).astype(np.int32)
    v0 = rng.uniform(-math.pi, math.pi, n - 3)
    ks = rng.uniform(-1.0, 1.0, n - 3) * KJ_PER_KCAL
    ns = rng.integers(1, 4, n - 3).astype(float)
    sysm.add(ForceSpec(abi.COSINE_DIHEDRAL, "CosineDihedral", indices=quad,
                       params=np.stack([ks, v0 + math.pi, ns], 1)))
    return sysm, integ, x
