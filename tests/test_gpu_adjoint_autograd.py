"""mps_hip_step_backward, the backward of one timestep, wrapped in a
torch.autograd.Function whose forward is the drop-in simulation_step.

Chained over steps of the 2-D dam break, autograd's gradient of the
position loss must equal mps_hip_adjoint's (the same reverse kernels; the
parameter gradients are summed per step on the host instead of on the
device).  The input-position adjoint, which mps_hip_adjoint does not
return, is checked against central differences of one step taken with the
same stale list.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import scenes
from mps_amd import abi as A
from oracle.oracle import make_config

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Sim:
    """The scene between steps: the host particle array and the stale
    neighbour list, as simulation_step passes them from step to step."""

    def __init__(self, lib, name="dam2d_l0.01", seed=7):
        self.lib = lib
        sc = scenes.CONFIGS[name]
        self.dim = sc["dim"]
        self.params = scenes.params_for(name, cg_tolerance=1e-13, cg_max_iter=20000, viscosity=1e-6)
        pos, vel, typ = scenes.dambreak(**sc)
        fl = typ == 0
        u0 = np.zeros_like(vel)
        u0[fl] = np.random.default_rng(seed).normal(0.0, 0.01, size=(fl.sum(), self.dim))
        lib.set_option("tol_mode", 1)
        cfg = self.cfg(1e-6, [0.0, -9.81])
        lib.set_config(cfg)
        s = A.HostScene(lib.abi, pos, u0, typ)
        lib.particle_system_calc_initial_params(s.ptr())
        self.n0, self.lam = s.ps.n0, s.ps.lambda_
        self.n = s.num
        re = max(cfg.influence_radius_lap, cfg.influence_radius_n)
        dmin = (C.c_double * self.dim)(*[cfg.domain_min[d] for d in range(self.dim)])
        dmax = (C.c_double * self.dim)(*[cfg.domain_max[d] for d in range(self.dim)])
        self.nl = lib.neighbor_list_create(self.n, cfg.max_neighbors)
        self.cl = lib.cell_list_create(self.n, re, dmin, dmax)
        lib.neighbor_search_cell_linked_list(self.nl, s.ptr(), self.cl, re)
        lib.mps_hip_release(s.ptr())
        self.state0 = (s.snapshot(), self.save_list())
        self.state = self.state0

    def cfg(self, nu, g):
        c = make_config(self.lib.abi, **dict(self.params, viscosity=float(nu)))
        for d in range(self.dim):
            c.gravity[d] = float(g[d])
        return c

    def save_list(self):
        m = self.nl.contents.max_neighbors
        cnt = np.ctypeslib.as_array(self.nl.contents.count, shape=(self.n,)).copy()
        nbr = np.ctypeslib.as_array(self.nl.contents.neighbors, shape=(self.n * m,)).copy()
        return cnt, nbr

    def load_list(self, lst):
        m = self.nl.contents.max_neighbors
        np.ctypeslib.as_array(self.nl.contents.count, shape=(self.n,))[:] = lst[0]
        np.ctypeslib.as_array(self.nl.contents.neighbors, shape=(self.n * m,))[:] = lst[1]

    def scene(self, arr, r=None, u=None):
        s = A.HostScene(self.lib.abi, arr["pos"], arr["vel"], arr["type"])
        s.arr[:] = arr
        if r is not None:
            s.arr["pos"] = r
        if u is not None:
            s.arr["vel"] = u
        s.ps.n0, s.ps.lambda_ = self.n0, self.lam
        return s

    def close(self):
        self.lib.neighbor_list_free(self.nl)
        self.lib.cell_list_free(self.cl)


class MPSStep(torch.autograd.Function):
    """(r, u) -> (r', u') of one simulation_step with viscosity nu and
    gravity g; the stale list and the rest of the particle state come from
    sim.state, which the forward advances."""

    @staticmethod
    def forward(ctx, r, u, nu, g, sim):
        arr, lst = sim.state
        ctx.sim, ctx.arr, ctx.lst = sim, arr.copy(), lst
        ctx.arr["pos"] = r.detach().numpy()
        ctx.arr["vel"] = u.detach().numpy()
        ctx.nu, ctx.g = float(nu), [float(x) for x in g]
        lib = sim.lib
        lib.set_config(sim.cfg(ctx.nu, ctx.g))
        s = sim.scene(ctx.arr)
        sim.load_list(lst)
        lib.simulation_step(s.ptr(), sim.nl, sim.cl, 1)
        assert lib.mps_hip_last_error() == 0
        lib.mps_hip_release(s.ptr())
        out = s.snapshot()
        sim.state = (out, sim.save_list())
        return torch.from_numpy(np.array(out["pos"])), torch.from_numpy(np.array(out["vel"]))

    @staticmethod
    def backward(ctx, rbar, ubar):
        sim, lib = ctx.sim, ctx.sim.lib
        lib.set_config(sim.cfg(ctx.nu, ctx.g))
        s = sim.scene(ctx.arr)
        sim.load_list(ctx.lst)
        rb = np.ascontiguousarray(rbar.numpy(), dtype=np.float64)
        ub = np.ascontiguousarray(ubar.numpy(), dtype=np.float64)
        rin, uin = np.zeros_like(rb), np.zeros_like(ub)
        gnu, gg = np.zeros(1), np.zeros(sim.dim)
        lib.check(lib.mps_hip_step_backward(s.ptr(), sim.nl, _p(rb), _p(ub), _p(rin), _p(uin), _p(gnu), _p(gg)),
                  "step_backward")
        return torch.from_numpy(rin), torch.from_numpy(uin), torch.tensor(gnu[0]), torch.from_numpy(gg), None


def _loss(r, typ, target):
    fl = torch.from_numpy(typ == 0)
    return ((r[fl] - target[fl]) ** 2).sum() / fl.sum()


def test_autograd_chain_equals_adjoint(lib2):
    steps = 6
    sim = Sim(lib2)
    arr0 = sim.state0[0]
    target = np.array(arr0["pos"]) + np.random.default_rng(8).normal(0.0, 1e-3, size=arr0["pos"].shape)
    r = torch.from_numpy(np.array(arr0["pos"]))
    u0 = torch.from_numpy(np.array(arr0["vel"])).requires_grad_(True)
    nu = torch.tensor(1e-6, dtype=torch.float64, requires_grad=True)
    g = torch.tensor([0.0, -9.81], dtype=torch.float64, requires_grad=True)
    u = u0
    for _ in range(steps):
        r, u = MPSStep.apply(r, u, nu, g, sim)
    typ = np.array(sim.state[0]["type"])
    loss = _loss(r, typ, torch.from_numpy(target))
    gu, gnu, gg = torch.autograd.grad(loss, [u0, nu, g])
    # the same gradient through mps_hip_adjoint
    lib2.set_config(sim.cfg(1e-6, [0.0, -9.81]))
    s = sim.scene(arr0)
    tgt = np.ascontiguousarray(target)
    L, anu, ag, au = np.zeros(1), np.zeros(1), np.zeros(2), np.zeros((sim.n, 2))
    lib2.check(lib2.mps_hip_adjoint(s.ptr(), steps, 4, _p(tgt), _p(L), _p(anu), _p(ag), _p(au)), "adjoint")
    sim.close()
    print("loss autograd %.15e adjoint %.15e" % (float(loss), L[0]), flush=True)
    assert float(loss) == pytest.approx(L[0], rel=1e-13)
    du = np.linalg.norm(gu.numpy() - au) / np.linalg.norm(au)
    print("u0 gradient rel diff %.3e, nu %.3e %.3e, g %s %s" % (du, float(gnu), anu[0], gg.numpy(), ag), flush=True)
    assert du < 1e-12
    assert float(gnu) == pytest.approx(anu[0], rel=1e-12)
    assert np.allclose(gg.numpy(), ag, rtol=1e-12, atol=0)


def test_position_adjoint_vs_finite_differences(lib2):
    sim = Sim(lib2)
    arr0 = sim.state0[0]
    fl = arr0["type"] == 0
    target = np.array(arr0["pos"]) + np.random.default_rng(9).normal(0.0, 1e-3, size=arr0["pos"].shape)
    tt = torch.from_numpy(target)
    nu = torch.tensor(1e-6, dtype=torch.float64)
    g = torch.tensor([0.0, -9.81], dtype=torch.float64)

    def one(r0, u0):
        sim.state = sim.state0
        r1, _ = MPSStep.apply(r0, u0, nu, g, sim)
        return _loss(r1, np.array(sim.state[0]["type"]), tt)

    r0 = torch.from_numpy(np.array(arr0["pos"])).requires_grad_(True)
    u0 = torch.from_numpy(np.array(arr0["vel"]))
    (gr,) = torch.autograd.grad(one(r0, u0), [r0])
    v = np.zeros_like(target)
    v[fl] = np.random.default_rng(10).normal(size=(fl.sum(), 2))
    h = 1e-9 / np.abs(v).max()     # moves no fluid particle by more than 1e-9 m
    with torch.no_grad():
        lp = float(one(r0 + h * torch.from_numpy(v), u0))
        lm = float(one(r0 - h * torch.from_numpy(v), u0))
    sim.close()
    fd = (lp - lm) / (2 * h)
    ad = float((gr.numpy() * v).sum())
    print("position adjoint %.10e fd %.10e" % (ad, fd), flush=True)
    assert abs(ad - fd) / max(abs(ad), abs(fd)) < 1e-4
    assert np.all(gr.numpy()[~fl] == 0.0)
