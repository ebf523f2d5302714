"""Output files of a run and the event catalogue built from them.

* ``StateLog`` writes ``state.txt`` in the reference driver's formats
  (src/main_gpu.py:49-53 header, :89 column line, :136 GPU-branch rows,
  :214 CPU-branch rows, end lines after the loop).
* ``write_legacy_vtk`` / ``read_legacy_vtk``: the legacy-VTK fault snapshot
  that ``QDsim.ouputVTK`` writes (layout of src/QDsim_gpu.py:1927-2073), and
  a reader of it.
* ``read_state`` parses ``state.txt`` rows.
* ``extract_slip_info`` restates the event catalogue of
  utils/Plot_tool.py:421-560 (``Ptool.extract_slip_info``) over ``state.txt``
  and the ``out_vtk/step<N>.vtk`` snapshots of a run directory, without
  pyvista: excursions of the logged maximum slip rate above Vdyn define the
  events; per event the snapshots inside it give the nucleation point, the
  extents of the cells above Vdyn, mean and maximum slip, the mean state and
  shear-stress changes, the seismic moment M0 (Simpson's rule over the moment
  rate G * sum(area * V)), the mean moment rate and Mw.

Host-side post-processing only; nothing here touches the GPU.
"""
import os
import re
from datetime import datetime

import numpy as np

# ----------------------------------------------------------------- state.txt


class StateLog:
    """``state.txt`` as the reference driver writes it.

    ``StateLog(path, fnamegeo, fnamePara)`` writes the three start lines
    (main_gpu.py:49-53); ``columns()`` the column line (:89); ``row_gpu`` /
    ``row_cpu`` one row per step of the GPU branch (:136) or the CPU branch
    (:214, which also logs Relerrormax1 / Relerrormax2); ``close(seconds)`` the
    two end lines."""

    def __init__(self, path, fnamegeo, fnamePara):
        self.f = open(path, "w")
        self.f.write("Program start time: %s\n" % str(datetime.now()))
        self.f.write("Input msh geometry file:%s\n" % fnamegeo)
        self.f.write("Input parameter file:%s\n" % fnamePara)

    def columns(self):
        self.f.write("iteration time_step(s) maximum_slip_rate(m/s) time(s) time(h)\n")

    def row_gpu(self, i, dttry, sim):
        """GPU branch: step, trial step, max |slipv1|, max |slipv2|, time (s),
        time (days) -- from the host copies the driver refreshes."""
        self.f.write("%d %f %f %.16e %f %f\n" % (i, dttry, np.max(np.abs(sim.slipv1)), np.max(np.abs(sim.slipv2)),
                                                 sim.time, sim.time / 3600.0 / 24.0))

    def row_cpu(self, i, dttry, sim):
        """CPU branch: step, trial step, max |slipv|, time (s), time (days),
        Relerrormax1, Relerrormax2."""
        self.f.write("%d %f %.16e %f %f %f %f\n" % (i, dttry, np.max(np.abs(sim.slipv)), sim.time,
                                                    sim.time / 3600.0 / 24.0, sim.Relerrormax1, sim.Relerrormax2))

    def close(self, seconds):
        self.f.write("Program end time: %s\n" % str(datetime.now()))
        self.f.write("Time taken: %.2f seconds\n" % seconds)
        self.f.close()


def read_state(path):
    """Rows of ``state.txt``: a float array with one row per step line (lines
    that are not all numbers -- the start, column and end lines -- are
    skipped).  GPU-branch rows have 6 columns (Iteration, dt, max |slipv1|,
    max |slipv2|, time (s), time (days)), CPU-branch rows 7."""
    rows = []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if len(tok) < 6:
                continue
            try:
                rows.append([float(t) for t in tok])
            except ValueError:
                continue
    if not rows:
        return np.zeros((0, 6))
    width = min(len(r) for r in rows)
    return np.array([r[:width] for r in rows])


# ----------------------------------------------------------------- legacy VTK


def write_legacy_vtk(fname, nodes, tri0, cell_fields):
    """Unstructured-grid snapshot of the fault: nodes (n, 3) as '%f', 0-based
    triangles, then each (label, values, number format) as a cell scalar in
    the given order (the layout of QDsim_gpu.py:1927-2073)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    tri0 = np.asarray(tri0, dtype=np.int64)
    ne = len(tri0)
    with open(fname, "w") as f:
        f.write("# vtk DataFile Version 3.0\ntest\nASCII\nDATASET  UNSTRUCTURED_GRID\n")
        f.write("POINTS %d float\n" % len(nodes))
        f.write("".join("%f %f %f\n" % tuple(p) for p in nodes))
        f.write("CELLS %d %d\n" % (ne, ne * 4))
        f.write("".join("3 %d %d %d\n" % tuple(t) for t in tri0))
        f.write("CELL_TYPES %d\n" % ne)
        f.write("5 " * ne + "\n")
        f.write("CELL_DATA %d " % ne)
        for label, vals, fmt in cell_fields:
            vals = np.asarray(vals, dtype=np.float64)
            f.write("SCALARS %s float\nLOOKUP_TABLE default\n" % label)
            f.write("".join((fmt + " ") % v for v in vals) + "\n")


def read_legacy_vtk(fname):
    """(points (n, 3), triangles (m, 3) 0-based, {label: cell values}) of a
    snapshot written by ``write_legacy_vtk``."""
    with open(fname) as f:
        tok = f.read().split()
    i = tok.index("POINTS")
    npts = int(tok[i + 1])
    pts = np.array(tok[i + 3:i + 3 + 3 * npts], dtype=np.float64).reshape(npts, 3)
    i = tok.index("CELLS")
    ncell, nint = int(tok[i + 1]), int(tok[i + 2])
    cells = np.array(tok[i + 3:i + 3 + nint], dtype=np.int64).reshape(ncell, 4)[:, 1:]
    data = {}
    i = tok.index("CELL_DATA") + 2
    while i < len(tok):
        if tok[i] != "SCALARS":
            i += 1
            continue
        label = tok[i + 1]
        j = i + 3
        if tok[j] == "LOOKUP_TABLE":
            j += 2
        data[label] = np.array(tok[j:j + ncell], dtype=np.float64)
        i = j + ncell
    return pts, cells, data


def _triangle_geometry(pts, tri0):
    p = pts[tri0]
    centroid = p.mean(axis=1)
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    return centroid, area


# ----------------------------------------------------------------- event catalogue

_HEADER = (f'{"Evnt":5}{"Nuc_X":10}{"Nuc_y":10}{"Nuc_z":10}{"X_min":10}{"X_max":10}{"Y_min":10}{"Y_max":10}'
           f'{"Z_min":10}{"Z_max":10}{"slip_mean":10}{"slip_max":10}{"State_drop":16}{"Stress_drop":16}'
           f'{"M0":16}{"M0_dot_mean":16}{"Mw":16}\n')


def excursions(state_rows, Vdyn=1e-3):
    """Runs of consecutive state.txt rows whose max slip rate (the norm of the
    logged strike and dip maxima, columns 2 and 3 of a GPU-branch row) exceeds
    Vdyn, as (first Iteration, last Iteration) pairs in file order."""
    r = np.asarray(state_rows, dtype=np.float64)
    if len(r) == 0:
        return []
    v = np.sqrt(r[:, 2] ** 2 + r[:, 3] ** 2)
    idx = np.nonzero(v > Vdyn)[0]
    if len(idx) == 0:
        return []
    breaks = np.nonzero(np.diff(idx) > 1)[0]
    starts = np.concatenate([[0], breaks + 1])
    ends = np.concatenate([breaks, [len(idx) - 1]])
    return [(int(r[idx[a], 0]), int(r[idx[b], 0])) for a, b in zip(starts, ends)]


def snapshot_steps(out_folder):
    """Iteration numbers of the ``step<N>.vtk`` / ``.vtu`` snapshots, sorted."""
    steps = []
    for name in os.listdir(out_folder):
        m = re.fullmatch(r"step(\d+)\.vt[ku]", name)
        if m:
            steps.append(int(m.group(1)))
    return np.sort(np.array(steps, dtype=np.int64))


def extract_slip_info(path, G=32038120320, Vdyn=1e-3, skip_ends=True, write=True):
    """Event catalogue of a run directory (``state.txt`` + ``out_vtk/``).

    Follows utils/Plot_tool.py:421-560:

    * events are the excursions of the logged max slip rate above Vdyn;
      with ``skip_ends`` (the reference's loop, ``range(Nevents - 1)`` over the
      starts after the first) the first and the last excursion are left out;
    * an event's snapshots are those whose iteration lies in its excursion;
      each gives the moment rate G * sum(area * V) and the extents of the
      cells with V > Vdyn; the first snapshot gives the nucleation point
      (centre of those extents) and the initial slip, shear stress and state,
      the last one the final values;
    * M0 = Simpson's rule of the moment rate over the snapshot times,
      Mw = 2/3 (log10 M0 - 9.1), M0_dot_mean = 10 ** mean(log10 moment rate);
      slip mean / max, state drop and stress drop are over all cells;
    * an event that raises (fewer than two snapshots, or a snapshot with no
      cell above Vdyn) is skipped, as the reference's try/except does.

    The slip field is read as 'slip[m]' (the reference's reader) or 'slip'
    (the label this build's and the reference's QDsim_gpu VTK writer use).
    Returns (list of per-event dicts, the events.txt text); the text is also
    written to ``path/events.txt`` unless ``write`` is false.  Parity with the
    pyvista-based reference is unpinned (pyvista is not available to run it):
    the restatement is checked against closed forms on constructed runs."""
    from scipy.integrate import simpson
    rows = read_state(os.path.join(path, "state.txt"))
    out_folder = os.path.join(path, "out_vtk")
    steps = snapshot_steps(out_folder)
    exc = excursions(rows, Vdyn)
    chosen = exc[1:-1] if skip_ends else exc
    text = _HEADER
    events = []
    for i, (it1, it2) in enumerate(chosen):
        sel = steps[(steps >= it1) & (steps <= it2)]
        try:
            if len(sel) < 2:
                raise ValueError("fewer than two snapshots inside the event")
            mo_dot = np.empty(len(sel))
            t = np.empty(len(sel))
            lo = np.full(3, np.inf)
            hi = np.full(3, -np.inf)
            for k, step in enumerate(sel):
                t[k] = rows[rows[:, 0] == step, 4][0]
                fn = os.path.join(out_folder, f"step{step}.vtu")
                if not os.path.exists(fn):
                    fn = os.path.join(out_folder, f"step{step}.vtk")
                pts, tri0, data = read_legacy_vtk(fn)
                cen, area = _triangle_geometry(pts, tri0)
                V = data["Slipv[m/s]"]
                mo_dot[k] = np.sum(area * V * G)
                fast = V > Vdyn
                if not np.any(fast):
                    raise ValueError(f"no cell above Vdyn in snapshot {step}")
                lo1, hi1 = cen[fast].min(axis=0), cen[fast].max(axis=0)
                lo, hi = np.minimum(lo, lo1), np.maximum(hi, hi1)
                slip = data["slip[m]"] if "slip[m]" in data else data["slip"]
                if k == 0:
                    nuc = 0.5 * (lo1 + hi1)
                    slip_ini, shear_ini, state_ini = slip, data["Shear_[MPa]"], data["state"]
                elif k == len(sel) - 1:
                    slip_end, shear_end, state_end = slip, data["Shear_[MPa]"], data["state"]
            M0 = simpson(mo_dot, x=t)
            ev = {"event": i, "first_iteration": it1, "last_iteration": it2, "nucleation": nuc,
                  "xyz_min": lo, "xyz_max": hi, "slip_mean": float((slip_end - slip_ini).mean()),
                  "slip_max": float((slip_end - slip_ini).max()),
                  "state_drop": float((state_end - state_ini).mean()),
                  "stress_drop": float((shear_end - shear_ini).mean()), "M0": float(M0),
                  "M0_dot_mean": float(10 ** np.mean(np.log10(mo_dot))),
                  "Mw": float(2.0 / 3.0 * (np.log10(M0) - 9.1))}
        except (ValueError, KeyError, IndexError) as exc_:
            print(f"event {i + 1}: skipped ({exc_})")
            continue
        events.append(ev)
        text += (f'{i:5.0f}{nuc[0]:10.1f}{nuc[1]:10.1f}{nuc[2]:10.1f}{lo[0]:10.1f}{hi[0]:10.1f}{lo[1]:10.1f}'
                 f'{hi[1]:10.1f}{lo[2]:10.1f}{hi[2]:10.1f}{ev["slip_mean"]:10.3f}{ev["slip_max"]:10.3f}'
                 f'{ev["state_drop"]:16.6E}{ev["stress_drop"]:16.6E}{ev["M0"]:16.6E}{ev["M0_dot_mean"]:16.6E}'
                 f'{ev["Mw"]:16.3f}\n')
    if write:
        with open(os.path.join(path, "events.txt"), "w") as f:
            f.write(text)
    return events, text
