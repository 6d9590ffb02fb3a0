"""A TOML input in the reference's format that uses every membrane and table
potential of the MD step: per-bead WCA, iSoLF attraction, table WCA and
table LJ attraction between named bead kinds, plus a capped Go contact,
a cosine dihedral and a Gaussian+Cosine dihedral on a short chain.

Lipids are three beads (H, T, T) in a periodic box, on a jittered lattice;
the chain is 6 beads along x.  Sizes follow the reference's POPC comments
(ReadTOMLForceFieldGenerator.cpp:1225-1236, :1423-1432): sigma_H 4.62 A,
sigma_T 7.11 A, epsilon 0.416 kcal/mol, omega 9.87 A.
"""
import numpy as np

BOX_A = 60.0


def positions(n_lipids=27, seed=0):
    rng = np.random.default_rng(seed)
    per = int(round(n_lipids ** (1.0 / 3.0)))
    grid = np.stack(np.meshgrid(*[np.arange(per)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n_lipids]
    spacing = BOX_A / per
    centres = (grid + 0.5) * spacing + rng.uniform(-1.0, 1.0, size=grid.shape)
    d = rng.normal(size=(n_lipids, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    lip = (centres[:, None, :] + d[:, None, :] * np.array([-5.5, 0.0, 5.5])[None, :, None]).reshape(-1, 3)
    chain = np.array([[3.0 + 3.8 * k, 3.0 + 0.9 * (k % 2), 1.5 + 0.7 * k] for k in range(6)])
    return np.concatenate([lip, chain]) % BOX_A


def toml_text(n_lipids=27, seed=0, capping_ratio=0.6, lj_cutoff=2.5):
    x = positions(n_lipids, seed)
    n_lip = 3 * n_lipids
    names = ["H", "T", "T"] * n_lipids + ["C"] * 6
    parts = ",\n".join(
        f'{{m = 100.0, pos = [{p[0]:.6f}, {p[1]:.6f}, {p[2]:.6f}], name = "{nm}"}}' for p, nm in zip(x, names))
    bonds = ",\n".join(f"{{indices = [{3 * k + a}, {3 * k + a + 1}], v0 = 5.5, k = 20.0}}"
                       for k in range(n_lipids) for a in (0, 1))
    chain_bonds = ",\n".join(f"{{indices = [{n_lip + k}, {n_lip + k + 1}], v0 = 3.8, k = 110.4}}" for k in range(5))
    angles = ",\n".join(f"{{indices = [{3 * k}, {3 * k + 1}, {3 * k + 2}], v0 = 3.141592653589793, k = 2.0}}"
                        for k in range(n_lipids))
    wca = ",\n".join(f'{{index = {i}, sigma = "{"sH" if i % 3 == 0 else "sT"}", epsilon = "eps"}}'
                     for i in range(n_lip))
    isolf = ",\n".join(f'{{index = {i}, sigma = "sT", epsilon = "eps", omega = "om"}}'
                       for i in range(n_lip) if i % 3 != 0)
    named = ",\n".join(f'{{index = {i}, name = "{names[i]}"}}' for i in range(n_lip))
    c = n_lip
    return f"""
[files]
output.prefix = "membrane"
output.format = "dcd"
[simulator]
boundary_type = "PeriodicCuboid"
delta_t = 0.2
total_step = 1000
save_step = 100
seed = 5
[[systems]]
boundary_shape.lower = [0.0, 0.0, 0.0]
boundary_shape.upper = [{BOX_A}, {BOX_A}, {BOX_A}]
attributes.temperature = 300.0
particles = [
{parts}
]
[[forcefields]]
[[forcefields.local]]
interaction = "BondLength"
potential = "Harmonic"
topology = "bond"
parameters = [
{bonds},
{chain_bonds}
]
[[forcefields.local]]
interaction = "BondAngle"
potential = "Harmonic"
topology = "angle"
parameters = [
{angles}
]
[[forcefields.local]]
interaction = "BondLength"
potential = "CappedGoContact"
topology = "contact"
capping_ratio = {capping_ratio}
parameters = [
{{indices = [{c}, {c + 4}], k = 0.3, v0 = 20.0}},
{{indices = [{c + 1}, {c + 5}], k = 0.5, v0 = 11.0}},
]
[[forcefields.local]]
interaction = "DihedralAngle"
potential = "Cosine"
topology = "none"
parameters = [
{{indices = [{c}, {c + 1}, {c + 2}, {c + 3}], v0 = -1.0, k = -0.4, n = 1}},
{{indices = [{c + 1}, {c + 2}, {c + 3}, {c + 4}], v0 = 0.5, k = 0.3, n = 3}},
]
[[forcefields.local]]
interaction = "DihedralAngle"
potential = "Gaussian+Cosine"
topology = "none"
parameters = [
{{indices = [{c + 2}, {c + 3}, {c + 4}, {c + 5}], Gaussian = {{v0 = 0.6, k = -1.0, sigma = 0.3}}, Cosine = {{v0 = -0.2, k = 0.5, n = 2}}}},
]
[[forcefields.global]]
potential = "WCA"
ignore.particles_within = {{bond = 1, angle = 1}}
env.eps = 0.416
env.sT = 7.111
env.sH = 4.62215
parameters = [
{wca}
]
[[forcefields.global]]
potential = "iSoLFAttractive"
ignore.particles_within = {{bond = 1, angle = 1}}
env.eps = 0.416
env.om = 9.867
env.sT = 7.111
parameters = [
{isolf}
]
[[forcefields.global]]
interaction = "Pair"
potential = "WCA"
ignore.particles_within = {{bond = 1}}
table.H.T = {{sigma = 5.5, epsilon = 0.3}}
table.C.H = {{sigma = 4.0, epsilon = 0.2}}
parameters = [
{named}
]
[[forcefields.global]]
interaction = "Pair"
potential = "LennardJonesAttractive"
cutoff = {lj_cutoff}
ignore.particles_within = {{bond = 1}}
table.T.T = {{sigma = 6.0, epsilon = 0.2}}
parameters = [
{named}
]
"""
