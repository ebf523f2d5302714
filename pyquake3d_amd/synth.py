"""Synthetic fault meshes and friction parameters for the benchmark workloads.

The workloads are written in the reference's own input formats -- gmsh 2.2
ASCII meshes in km (read by ``read_mshV2``, src/readmsh.py:12-68), the
10-column per-element init file (``read_parameter``, src/QDsim_gpu.py:714-770)
and a ``key: value`` parameter file (src/QDsim_gpu.py:90-99) -- so the
reference path can ingest them unchanged.

Workloads (sizes as in the benchmark definition):

* ``planar``: vertical strike-slip fault in the plane y = 0, nx x nz squares
  each split along the same diagonal, top edge at the free surface.
    - 80 km x 40 km, 40 x 20 squares  -> 1,600 triangles
    - 40 km x 40 km, 100 x 100 squares -> 20,000 triangles
* friction: aging law, b = 0.03, Dc = 0.14 m, a = 0.004 in a central
  30 km x 15 km velocity-weakening patch, rising linearly to 0.04 over a
  5 km taper outside it; normal stress 25 MPa; V = Vpl = 1e-9 m/s with the
  state at steady state; optional seeded lognormal Dc perturbation;
* ensemble members: a - b ~ U[-0.006, -0.002] on the patch,
  Dc = 0.14 exp(0.3 N(0,1)), sigma_n ~ U[20, 30] MPa, seed = base + member.
"""
import os

import numpy as np

MU = 32038120320.0
LAMBDA = 32038120320.0
CS = 3464.0
RHO = MU / CS ** 2
V0 = 1e-6
F0 = 0.6
VPL = 1e-9
SIGMA_N = 25.0  # MPa
TAPER_KM = 5.0


def planar_fault_mesh(nx, nz, length_km, depth_km, x0_km=None):
    """Nodes (km) and 1-based triangles of a vertical fault in y = 0."""
    if x0_km is None:
        x0_km = -0.5 * length_km
    xs = x0_km + length_km * np.arange(nx + 1) / nx
    zs = -depth_km + depth_km * np.arange(nz + 1) / nz
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    nodes = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], axis=1)

    def nid(i, k):
        return i * (nz + 1) + k + 1

    tri = []
    for i in range(nx):
        for k in range(nz):
            a, b, c, d = nid(i, k), nid(i + 1, k), nid(i + 1, k + 1), nid(i, k + 1)
            # both halves ordered so that (b - a) x (c - a) points along +y
            tri.append((a, c, b))
            tri.append((a, d, c))
    return nodes, np.array(tri, dtype=np.int64)


def rough_fault_mesh(nx=300, nz=100, length_km=120.0, depth_km=30.0, amp_km=1.0, wavelength_km=40.0,
                     hurst=0.8, rough_ratio=1e-3, seed=7):
    """Nodes (km) and 1-based triangles of a non-planar vertical fault.

    Near-equilateral lattice: nz + 1 node rows from z = -depth to 0, each with
    nx + 1 nodes spaced length/nx, odd rows shifted by half a spacing; every
    strip between two rows holds 2 nx triangles, 2 nx nz in all (300 x 100 ->
    60,000 triangles of sides 400 m and about 360 m).  The fault-normal offset
    is y = amp sin(2 pi x / wavelength) + h(x, z), with h a seeded self-affine
    random surface (Fourier amplitudes ~ |k|^-(1 + hurst), random phases)
    scaled to an RMS height of rough_ratio x length.  Triangles are ordered so
    that their normals point along +y.
    """
    dx = length_km / nx
    ks = np.arange(nz + 1)
    js = np.arange(nx + 1)
    X = -0.5 * length_km + dx * (js[None, :] + 0.5 * (ks[:, None] % 2))
    Z = np.broadcast_to((-depth_km + depth_km * ks / nz)[:, None], X.shape)
    rng = np.random.default_rng(seed)
    kz = np.fft.fftfreq(nz + 1)[:, None]
    kx = np.fft.fftfreq(nx + 1)[None, :]
    kk = np.hypot(kx, kz)
    kk[0, 0] = np.inf
    spec = kk ** (-(1.0 + hurst)) * np.exp(2j * np.pi * rng.random(kk.shape))
    h = np.fft.ifft2(spec).real
    h *= rough_ratio * length_km / np.sqrt(np.mean(h * h))
    Y = amp_km * np.sin(2.0 * np.pi * X / wavelength_km) + h
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def nid(k, j):
        return k * (nx + 1) + j + 1

    tri = []
    for k in range(nz):
        for j in range(nx):
            if k % 2 == 0:   # upper row shifted right
                tri.append((nid(k, j), nid(k, j + 1), nid(k + 1, j)))
                tri.append((nid(k + 1, j), nid(k, j + 1), nid(k + 1, j + 1)))
            else:            # lower row shifted right
                tri.append((nid(k, j), nid(k + 1, j + 1), nid(k + 1, j)))
                tri.append((nid(k, j), nid(k, j + 1), nid(k + 1, j + 1)))
    tri = np.array(tri, dtype=np.int64)
    p = nodes[tri - 1]
    nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    flip = nrm[:, 1] < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    return nodes, tri


def write_msh(fname, nodes_km, tri):
    with open(fname, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(nodes_km))
        for i, p in enumerate(nodes_km):
            f.write("%d %.12g %.12g %.12g\n" % (i + 1, p[0], p[1], p[2]))
        f.write("$EndNodes\n$Elements\n%d\n" % len(tri))
        for i, t in enumerate(tri):
            f.write("%d 2 2 0 1 %d %d %d\n" % (i + 1, t[0], t[1], t[2]))
        f.write("$EndElements\n")


def patch_distance_km(xg_m, half_len_km=15.0, zc_km=-20.0, half_h_km=7.5):
    """Euclidean distance (km) from element centroids to the central patch."""
    x = xg_m[:, 0] / 1e3
    z = xg_m[:, 2] / 1e3
    dx = np.maximum(np.abs(x) - half_len_km, 0.0)
    dz = np.maximum(np.abs(z - zc_km) - half_h_km, 0.0)
    return np.sqrt(dx * dx + dz * dz)


def friction_params(xg_m, dc_sigma=0.0, seed=0, a_vw=0.004, a_vs=0.04, b=0.03, dc=0.14, sigma_n=SIGMA_N,
                    rake=0.0):
    """Per-element (rake, a, b, Dc, f0, Tt, Tn, V, shear_loading, normal_loading)
    in the column order of the reference's init file."""
    n = len(xg_m)
    d = patch_distance_km(xg_m)
    a = a_vw + (a_vs - a_vw) * np.clip(d / TAPER_KM, 0.0, 1.0)
    dcv = np.full(n, dc)
    if dc_sigma > 0:
        rng = np.random.default_rng(seed)
        dcv = dc * np.exp(dc_sigma * rng.standard_normal(n))
    bb = np.full(n, b)
    sn = np.broadcast_to(np.asarray(sigma_n, dtype=np.float64), (n,)).copy()
    V = np.full(n, VPL)
    tt = sn * a * np.arcsinh(V / (2 * V0) * np.exp((F0 + bb * np.log(V0 / VPL)) / a))
    return np.stack([np.full(n, rake), a, bb, dcv, np.full(n, F0), tt, sn, V, np.zeros(n), np.zeros(n)], axis=1)


def member_params(xg_m, member, base_seed=1000, amb_range=(-0.006, -0.002), dc_sigma=0.3, rake=0.0):
    """One ensemble member's friction parameters (seed = base + member):
    a - b uniform in ``amb_range`` on the velocity-weakening patch, Dc
    lognormal per element (median 0.14 m, ``dc_sigma``), normal stress
    uniform 20-30 MPa, slip direction ``rake`` (radians from strike).
    With the default range the critical nucleation size (about
    mu Dc / ((b - a) sigma), 30-90 km) exceeds the 30 x 15 km patch, so the
    members creep without seismic events; a range near -0.024 puts it near
    7 km and the members rupture."""
    rng = np.random.default_rng(base_seed + member)
    n = len(xg_m)
    d = patch_distance_km(xg_m)
    b = np.full(n, 0.03)
    amb = rng.uniform(*amb_range)
    a_vw = 0.03 + amb
    a = a_vw + (0.04 - a_vw) * np.clip(d / TAPER_KM, 0.0, 1.0)
    dc = 0.14 * np.exp(dc_sigma * rng.standard_normal(n))
    sn = np.full(n, rng.uniform(20.0, 30.0))
    V = np.full(n, VPL)
    tt = sn * a * np.arcsinh(V / (2 * V0) * np.exp((F0 + b * np.log(V0 / VPL)) / a))
    return np.stack([np.full(n, rake), a, b, dc, np.full(n, F0), tt, sn, V, np.zeros(n), np.zeros(n)], axis=1)


def write_param_file(fname, cols):
    np.savetxt(fname, cols, fmt="%.8E")


PARAMETER_TXT = """#Files and cells
Corefunc directory: {core}
save Corefunc: True
Node_order: False
Scale_km: True

#Only availible for single GPU/CPU version
Max_thread workers: 8
Batch_size: 1000
GPU:{gpu}

#Property
Lame constants: {lam!r}
Shear modulus: {mu!r}
Rock density: {rho!r}
InputHetoparamter: True
Inputparamter file: {param}

#Fluid
If Dilatancy:False
Dilatancy coefficient: 3e-4
Hydraulic diffusivity: 1e-5
Actively shearing zone thickness: 3e-3
Effective compressibility: 8e-11
Constant porepressure: 2
Initial porepressure: 2

#Stress
Half space: {halfspace}
Fix_Tn:False
ssv_scale: 1.0
ssh1_scale: 1.6
ssh2_scale: 0.6
Angle between ssh1 and X-axis: 30
Vertical principal stress value: 25e6
Vertical principal stress value varies with depth: False
Turnning depth: 15000
Shear traction solved from stress tensor: False
Normal traction solved from stress tensor: False
Rake solved from stress tensor: False
Fix_rake: 0
Widths of VS region: 10000
Widths of surface VS region: 2000
Transition region from VS to VW region: 2000

#Friction
Reference slip rate: {v0!r}
Reference friction coefficient: {f0!r}
Plate loading rate: {vpl!r}

#output
totaloutputsteps: {steps}
outsteps: 50
outputvtu: True
outputSLIPV: False
"""


def write_case(dirname, nx, nz, length_km, depth_km, dc_sigma=0.0, seed=0, steps=2000, gpu=True,
               halfspace=True):
    """Write mesh.msh, param.dat and parameter.txt for one planar case."""
    from .readmsh import get_eleVec
    os.makedirs(dirname, exist_ok=True)
    nodes_km, tri = planar_fault_mesh(nx, nz, length_km, depth_km)
    write_msh(os.path.join(dirname, "mesh.msh"), nodes_km, tri)
    _, xg = get_eleVec(nodes_km * 1e3, tri, False)
    write_param_file(os.path.join(dirname, "param.dat"), friction_params(xg, dc_sigma, seed))
    with open(os.path.join(dirname, "parameter.txt"), "w") as f:
        f.write(PARAMETER_TXT.format(core=os.path.join(dirname, "core"), gpu=gpu, lam=LAMBDA, mu=MU, rho=RHO,
                                     param="param.dat", halfspace=halfspace, v0=V0, f0=F0, vpl=VPL,
                                     steps=steps))
    return os.path.join(dirname, "mesh.msh"), os.path.join(dirname, "parameter.txt")
