"""CPU tests of the run outputs: state.txt in the reference driver's formats
(main_gpu.py:49-53, :89, :136, :214), the legacy-VTK snapshot round trip, and
the event catalogue (utils/Plot_tool.py:421-560 restated in
pyquake3d_amd.outputs.extract_slip_info) on a constructed run whose answers
follow in closed form:

* four excursions of the logged max slip rate above Vdyn; the reference's
  loop leaves out the first and the last, so two events are catalogued;
* inside an event every snapshot has slip rate Vev on a patch of cells (a
  smaller nucleation patch in the first snapshot) and 1e-9 m/s elsewhere, so
  the moment rate G * sum(area_i V_i) takes two known values and M0 is
  composite Simpson over the five snapshots, Mw = 2/3 (log10 M0 - 9.1), the
  mean moment rate the geometric mean of the five;
* slip grows by V * 20 s between snapshots, so slip_max and slip_mean follow
  from the same rates; the shear stress and state change by given amounts on
  the patch;
* extents are the bounding box of the patch centroids, the nucleation point
  the centre of the nucleation patch's box;
* an event whose snapshots show no cell above Vdyn is skipped.

Parity with the pyvista-based reference itself is unpinned (pyvista is not
available to run it).
"""
import os

import numpy as np

from pyquake3d_amd import outputs, synth

G = 32038120320.0
VDYN = 1e-3


class _Sim:
    pass


def test_state_log_formats(tmp_path):
    p = str(tmp_path / "state.txt")
    log = outputs.StateLog(p, "mesh.msh", "parameter.txt")
    log.columns()
    s = _Sim()
    s.slipv1 = np.array([1e-9, -3.5e-2])
    s.slipv2 = np.array([2e-10, 1.25e-3])
    s.slipv = np.hypot(s.slipv1, s.slipv2)
    s.time = 86400.0 * 2.5
    s.Relerrormax1, s.Relerrormax2 = 3.3e-5, 8.2e-6
    log.row_gpu(7, 0.125, s)
    log.row_cpu(8, 0.25, s)
    log.close(12.5)
    lines = open(p).read().split("\n")
    assert lines[0].startswith("Program start time: ")
    assert lines[1] == "Input msh geometry file:mesh.msh"
    assert lines[2] == "Input parameter file:parameter.txt"
    assert lines[3] == "iteration time_step(s) maximum_slip_rate(m/s) time(s) time(h)"
    # GPU row: '%d %f %f %.16e %f %f' = step, dt, max|v1|, max|v2|, time, days
    assert lines[4] == "7 0.125000 0.035000 1.2500000000000000e-03 216000.000000 2.500000"
    # CPU row: '%d %f %.16e %f %f %f %f' = step, dt, max|v|, time, days, errors
    v = np.max(s.slipv)
    assert lines[5] == "8 0.250000 %.16e 216000.000000 2.500000 0.000033 0.000008" % v
    assert lines[6].startswith("Program end time: ") and lines[7] == "Time taken: 12.50 seconds"
    rows = outputs.read_state(p)
    assert rows.shape == (2, 6)       # the narrower row's width
    assert rows[0, 0] == 7 and rows[1, 0] == 8 and rows[0, 3] == 1.25e-3


def test_vtk_round_trip(tmp_path):
    nodes_km, tri = synth.planar_fault_mesh(6, 3, 12.0, 6.0)
    nodes = nodes_km * 1e3
    a = np.linspace(0.0, 1.0, len(tri))
    fn = str(tmp_path / "s.vtk")
    outputs.write_legacy_vtk(fn, nodes, tri - 1, [("Slipv[m/s]", a * 1e-3, "%.20f"), ("state", a, "%f")])
    pts, cells, data = outputs.read_legacy_vtk(fn)
    assert np.array_equal(cells, tri - 1)
    assert np.allclose(pts, nodes, atol=1e-6)
    assert list(data) == ["Slipv[m/s]", "state"]
    assert np.allclose(data["Slipv[m/s]"], a * 1e-3, rtol=0, atol=1e-19)
    assert np.allclose(data["state"], a, rtol=0, atol=1e-6)


def _write_run(path):
    """A constructed run: state.txt rows 0..99 (t = 10 s per step) with four
    excursions above Vdyn, and snapshots every 2 steps."""
    os.makedirs(os.path.join(path, "out_vtk"))
    nodes_km, tri = synth.planar_fault_mesh(10, 5, 20.0, 10.0)
    nodes = nodes_km * 1e3
    tri0 = tri - 1
    p = nodes[tri0]
    cen = p.mean(axis=1)
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    n = len(tri0)
    exc = [(10, 14), (30, 38), (60, 68), (85, 89)]
    vev = {1: 0.5, 2: 2.0}                     # slip rate on the patch in the catalogued events
    patch = (np.abs(cen[:, 0]) < 4000) & (cen[:, 2] > -7000) & (cen[:, 2] < -2000)
    nuc = patch & (cen[:, 0] > 0) & (cen[:, 0] < 4000) & (cen[:, 2] < -4000)
    assert nuc.sum() >= 2 and patch.sum() > nuc.sum()
    t = 10.0 * np.arange(100)
    log = outputs.StateLog(os.path.join(path, "state.txt"), "m", "p")
    log.columns()
    s = _Sim()
    for i in range(100):
        k = [j for j, (a, b) in enumerate(exc) if a <= i <= b]
        vmax = (vev.get(k[0], 0.01) if k else 1e-9)
        s.slipv1 = np.array([vmax])
        s.slipv2 = np.array([0.0])
        s.time = t[i]
        log.row_gpu(i, 10.0, s)
    log.close(1.0)
    slip = np.zeros(n)
    for i in range(0, 100, 2):
        k = [j for j, (a, b) in enumerate(exc) if a <= i <= b]
        V = np.full(n, 1e-9)
        shear = np.full(n, 10.0)
        state = np.full(n, 0.7)
        if k and k[0] in vev:
            a, b = exc[k[0]]
            V[nuc if i == a else patch] = vev[k[0]]
            if i == b:                       # end of the event: shear and state dropped on the patch
                shear[patch] -= 2.0 * k[0]
                state[patch] -= 0.1 * k[0]
        if k and k[0] == 3:
            V[:] = 1e-9                      # an excursion whose snapshots show no fast cell
        fields = [("Shear_[MPa]", shear, "%f"), ("state", state, "%f"), ("Slipv[m/s]", V, "%.20f"),
                  ("slip", slip, "%.20f")]
        outputs.write_legacy_vtk(os.path.join(path, "out_vtk", f"step{i}.vtk"), nodes, tri0, fields)
        slip = slip + V * 20.0               # slip grows until the next snapshot (2 steps of 10 s)
    return dict(exc=exc, vev=vev, patch=patch, nuc=nuc, cen=cen, area=area, t=t, n=n)


def test_event_catalogue_closed_forms(tmp_path):
    path = str(tmp_path)
    c = _write_run(path)
    assert outputs.excursions(outputs.read_state(os.path.join(path, "state.txt")), VDYN) == c["exc"]
    events, text = outputs.extract_slip_info(path, G=G, Vdyn=VDYN)
    assert open(os.path.join(path, "events.txt")).read() == text
    lines = text.split("\n")
    assert lines[0].split() == ["Evnt", "Nuc_X", "Nuc_y", "Nuc_z", "X_min", "X_max", "Y_min", "Y_max", "Z_min",
                                "Z_max", "slip_mean", "slip_max", "State_drop", "Stress_drop", "M0",
                                "M0_dot_mean", "Mw"]
    # the reference's loop catalogues excursions 1 and 2 (of 0..3); excursion 2's
    # snapshots are fine, excursion 3 is never considered
    assert [e["event"] for e in events] == [0, 1]
    cen, area, patch, nuc = c["cen"], c["area"], c["patch"], c["nuc"]
    for ev, k in zip(events, (1, 2)):
        a, b = c["exc"][k]
        vev = c["vev"][k]
        T = c["t"][b] - c["t"][a]
        # the moment rate at the five snapshots (only the nucleation patch is
        # fast in the first)
        ts = c["t"][a:b + 1:2]
        rate = np.full(len(ts), G * (np.sum(area[patch]) * vev + np.sum(area[~patch]) * 1e-9))
        rate[0] = G * (np.sum(area[nuc]) * vev + np.sum(area[~nuc]) * 1e-9)
        h = ts[1] - ts[0]
        M0 = h / 3.0 * (rate[0] + 4 * rate[1] + 2 * rate[2] + 4 * rate[3] + rate[4])   # composite Simpson
        assert abs(ev["M0"] - M0) <= 1e-12 * M0
        assert abs(ev["Mw"] - 2.0 / 3.0 * (np.log10(M0) - 9.1)) <= 1e-12
        assert abs(ev["M0_dot_mean"] - 10 ** np.mean(np.log10(rate))) <= 1e-12 * ev["M0_dot_mean"]
        # slip between the first and last snapshots: nucleation cells slip from
        # the first snapshot on, the rest of the patch from the second
        dslip = np.full(c["n"], 1e-9 * T)
        dslip[nuc] = vev * T
        dslip[patch & ~nuc] = 1e-9 * 20.0 + vev * (T - 20.0)
        assert abs(ev["slip_max"] - dslip.max()) <= 1e-9 * dslip.max()
        assert abs(ev["slip_mean"] - dslip.mean()) <= 1e-9 * dslip.mean()
        assert abs(ev["stress_drop"] - (-2.0 * k) * patch.mean()) <= 1e-6
        assert abs(ev["state_drop"] - (-0.1 * k) * patch.mean()) <= 1e-6
        assert np.allclose(ev["xyz_min"], cen[patch].min(axis=0), atol=1e-6)
        assert np.allclose(ev["xyz_max"], cen[patch].max(axis=0), atol=1e-6)
        assert np.allclose(ev["nucleation"], 0.5 * (cen[nuc].min(axis=0) + cen[nuc].max(axis=0)), atol=1e-6)
    row = lines[1]
    assert row[:5] == "    0" and len(row) == 5 + 9 * 10 + 2 * 10 + 5 * 16
    assert float(row[-16:]) == round(events[0]["Mw"], 3)


def test_event_catalogue_all_excursions(tmp_path):
    """skip_ends=False also considers the first and last excursions; the last
    one's snapshots have no fast cell and it is skipped, the first (its
    snapshots at the default 0.01 m/s on no cell) likewise."""
    path = str(tmp_path)
    _write_run(path)
    events, _ = outputs.extract_slip_info(path, G=G, Vdyn=VDYN, skip_ends=False, write=False)
    assert [(e["first_iteration"], e["last_iteration"]) for e in events] == [(30, 38), (60, 68)]
    assert not os.path.exists(os.path.join(path, "events.txt"))
