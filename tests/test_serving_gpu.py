"""GPU tests of the continuous-batching server.

* A request's result (simulated time, state, slip) is bit-identical whether it
  runs alone or admitted among others arriving at random times on two meshes.
* With an HBM budget that holds one mesh at a time, requests for both meshes
  all complete and the resident footprint (meshes resident or loading) never
  exceeds the budget; the budget formula bounds the memory the library
  reports for a resident mesh.
* Step graphs built and launched on one thread while another thread loads a
  mesh (allocations, copies, synchronous calls) come out whole.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

YEAR = 365.0 * 24 * 3600


def _meshes():
    from pyquake3d_amd import readmsh, synth
    out, xg = {}, {}
    for name, (nx, nz, L, D) in {"a": (10, 5, 20.0, 10.0), "b": (16, 8, 32.0, 16.0)}.items():
        nodes_km, tri = synth.planar_fault_mesh(nx, nz, L, D)
        nodes = nodes_km * 1e3
        out[name] = (nodes, tri)
        xg[name] = readmsh.get_eleVec(nodes, tri, False)[1]
    return out, xg


def _requests(xg, count, seed, years=(0.2, 1.0), rate=200.0):
    from pyquake3d_amd.serving import poisson_requests
    return poisson_requests(xg, rate, count, seed=seed, years=years)


def test_result_same_alone_and_in_the_mix():
    from pyquake3d_amd.serving import Server
    meshes, xg = _meshes()
    reqs = _requests(xg, 12, seed=5)
    probe = reqs[7]
    alone = Server(meshes, 64e9, steps_per_iteration=16, graph_steps=8)
    solo = type(probe)(rid=probe.rid, mesh=probe.mesh, cols=probe.cols, duration_s=probe.duration_s, arrival=0.0)
    m1 = alone.run([solo])
    alone.close()
    mixed = Server(meshes, 64e9, steps_per_iteration=16, graph_steps=8)
    m2 = mixed.run(reqs)
    mixed.close()
    assert m1["requests"] == 1 and m2["requests"] == len(reqs) and m2["failed"] == 0
    got = [r for r in mixed.done if r.rid == probe.rid][0]
    ref = alone.done[0]
    assert got.steps == ref.steps and got.result["time"] == ref.result["time"]
    for k in ("state", "slip", "slipv"):
        assert np.array_equal(got.result[k], ref.result[k]), k
    assert all(r.result["time"] >= r.duration_s for r in mixed.done)
    assert m2["p50_time_to_result_s"] <= m2["p99_time_to_result_s"]
    assert 0.0 < m2["mean_occupancy"] <= 1.0 and m2["member_steps"] > 0


def test_admission_under_hbm_budget():
    from pyquake3d_amd.serving import Server, _Mesh
    meshes, xg = _meshes()
    sizes = {k: _Mesh(k, *v, 1.0, 1.0, True).hbm_bytes() for k, v in meshes.items()}
    budget = max(sizes.values()) * 1.05   # one mesh at a time
    srv = Server(meshes, budget, steps_per_iteration=16, graph_steps=8)
    peak = [0]
    orig = srv._start_loads

    def start_loads():
        orig()
        peak[0] = max(peak[0], srv._resident_bytes())

    srv._start_loads = start_loads
    m = srv.run(_requests(xg, 8, seed=9))
    # the budget formula bounds what the library reports for a resident mesh
    for mesh in srv.meshes.values():
        if mesh.resident:
            held = mesh.sim.device_bytes() + mesh.ens.device_bytes()
            assert held <= mesh.hbm_bytes(), (mesh.name, held, mesh.hbm_bytes())
    srv.close()
    assert m["requests"] == 8 and m["failed"] == 0
    assert peak[0] <= budget


def test_graphs_built_while_another_thread_loads():
    import threading
    from pyquake3d_amd.serving import _Mesh
    meshes, _ = _meshes()
    mu = 32038120320.0
    a = _Mesh("a", *meshes["a"], mu, mu, True)
    a.load(tune=False)
    stop, errors, loads = threading.Event(), [], [0]

    def loader():
        try:
            while not stop.is_set():
                b = _Mesh("b", *meshes["b"], mu, mu, True)
                b.load(tune=False)
                b.unload()
                loads[0] += 1
        except Exception as e:  # reported below
            errors.append(e)

    t = threading.Thread(target=loader)
    t.start()
    k = 0
    try:
        while loads[0] < 3 and not errors:   # graphs built all through three loads
            g = 1 + k % 3            # a different graph length each call: a new graph every time
            a.ens.run(2 * g, graph_steps=g)
            k += 1
            assert k < 100000
    finally:
        stop.set()
        t.join()
    a.unload()
    assert not errors, errors
    print(f"{k} graphs built during {loads[0]} loads", flush=True)
