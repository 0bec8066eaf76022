"""Several grids advanced together with a nest (wam_gpu_multigrid_*), through
the C-ABI: configuration 4's three grids, shrunk (0.5 deg 24 x 25 beside a
0.2 deg -> 0.05 deg 36 x 36 nest).  Three steps grid after grid on one
stream and as one graph with the grids on concurrent branches must both
equal plain runs of each grid (wam_gpu_propags_dev, wam_gpu_implsch_dev)
with the nest exchange of the existing entry points in between
(wam_gpu_boundary_output_dev on the coarse grid, wam_gpu_boundary_input_dev
with the set at the model time into the fine grid), bit for bit; those
entry points are pinned against the oracle and the reference in
test_gpu_nest.py and test_ref_golden_nest.py."""
import numpy as np
import pytest

from wam_amd import WamError, WamGpu, WamMultiGrid, workloads as W

pytestmark = pytest.mark.gpu


def bits(t):
    return t.cpu().numpy().view(np.int32)


def make(torch, scale=0.08):
    grids, (c, f, link) = W.c4_case(scale=scale)
    dev = torch.device("cuda:0")
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    out = []
    for cfg, mask, fl in grids:
        g = WamGpu(cfg, mask)
        g.prepare_propagation(fl["depth"], fl["u"], fl["v"])
        n = g.nsea
        st = dict(fa=t(W.jonswap_spectrum(cfg, fl["udir"]).ravel()), u10=t(fl["u10"]), udir=t(fl["udir"]),
                  roairn=torch.full((n,), 1.225, device=dev), wstar=torch.zeros(n, device=dev),
                  tauw=torch.zeros(n, device=dev), ustar=torch.zeros(n, device=dev), z0=torch.zeros(n, device=dev))
        st["fb"] = torch.zeros_like(st["fa"])
        out.append((g, st))
    return out, (c, f, link)


def snapshot(grids):
    return [{k: v.clone() for k, v in st.items()} for _, st in grids]


def restore(grids, snap):
    for (_, st), s in zip(grids, snap):
        for k, v in s.items():
            st[k].copy_(v)


def plain(torch, grids, link, nsteps):
    """Each grid's steps with the plain entry points, then the nest exchange."""
    c, f, (ijarc, ijarf, ibfl, ibfr, bfw) = link
    cur = ["fa"] * len(grids)
    for _ in range(nsteps):
        for i, (g, st) in enumerate(grids):
            a, b = cur[i], "fb" if cur[i] == "fa" else "fa"
            g.propags_dev(st[a], st[b])
            g.implsch_dev(st[b], st["u10"], st["udir"], st["tauw"], st["ustar"], st["z0"], st["roairn"], st["wstar"])
            cur[i] = b
        gc, sc = grids[c]
        gf, sf = grids[f]
        fbc, mean = gc.boundary_output(sc[cur[c]], ijarc)
        gf.boundary_input(sf[cur[f]], fbc, mean, None, None, 1, 0.0, 0.0, ijarf, ibfl, ibfr, bfw, return_set=False)
    torch.cuda.synchronize()
    return [(st[cur[i]].clone(), st["tauw"].clone(), st["ustar"].clone(), st["z0"].clone())
            for i, (_, st) in enumerate(grids)]


def test_nested_runs_equal_plain_runs():
    import torch
    grids, (c, f, link) = make(torch)
    print(f"[case] NSEA {[g.nsea for g, _ in grids]}, nest {link[0].size} -> {link[1].size} points", flush=True)
    snap = snapshot(grids)
    want = plain(torch, grids, (c, f, link), 3)
    for fused in (False, True, True):       # the second fused run replays the graph
        restore(grids, snap)
        mg = WamMultiGrid(grids)
        mg.add_nest(c, f, *link)
        mg.run(3, fused=fused)
        torch.cuda.synchronize()
        for i, (g, st) in enumerate(grids):
            got = st[mg.state(i)]
            assert np.array_equal(bits(got), bits(want[i][0])), (fused, i)
            for j, k in enumerate(("tauw", "ustar", "z0")):
                assert np.array_equal(bits(st[k]), bits(want[i][1 + j])), (fused, i, k)
        mg.close()
    for g, _ in grids:
        g.close()


def test_graph_replays_continue_the_run():
    """Two fused runs of two steps equal one sequential run of four."""
    import torch
    grids, (c, f, link) = make(torch)
    snap = snapshot(grids)
    mg = WamMultiGrid(grids)
    mg.add_nest(c, f, *link)
    mg.run(4, fused=False)
    torch.cuda.synchronize()
    want = [st[mg.state(i)].clone() for i, (_, st) in enumerate(grids)]
    mg.close()
    restore(grids, snap)
    mg = WamMultiGrid(grids)
    mg.add_nest(c, f, *link)
    mg.run(2, fused=True)
    mg.run(2, fused=True)
    torch.cuda.synchronize()
    for i, (_, st) in enumerate(grids):
        assert np.array_equal(bits(st[mg.state(i)]), bits(want[i])), i
    mg.close()


def test_nest_needs_matching_spectra():
    import torch
    grids, (c, f, link) = make(torch)
    mg = WamMultiGrid(grids)
    with pytest.raises(WamError) as e:           # the 24 x 25 grid cannot feed a 36 x 36 one
        mg.add_nest(0, f, np.array([1], np.int32), np.array([1], np.int32), np.array([1], np.int32),
                    np.array([1], np.int32), np.array([0.0], np.float32))
    assert e.value.code == 1
    mg.close()
