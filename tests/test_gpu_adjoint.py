"""Gradients of the 2-D dam break through mps_hip_adjoint against central
finite differences of the GPU forward (relative step 1e-6).

Loss: mean squared distance of the fluid particles' final positions to a
target snapshot made by the same solver with nu = 5e-5, displaced by a
seeded 1 mm random offset -- without it the target lies so close to the
trajectory (loss 1e-15..1e-11) that the rounding of the positions, not the
derivative, sets the finite differences.  Parameters:
viscosity, gravity (both components) and the initial velocities (checked
along a random direction).  The reference has no derivatives, so finite
differences of the forward are the check -- of the GPU forward, and for
the longest case also of the CPU oracle's forward (the reference's
arithmetic restated).  The CG runs to a relative residual of 1e-13 so that
the differences resolve the loss.  Steps: 1e-6 relative for gravity and
the velocities; for the viscosity 1e-6 of the target's 5e-5 (1e-6 of the
base 1e-6 moves the loss by ~1e-14 relative, the loss's own rounding);
after a single step, where that step moves the loss by ~1e-16 relative,
1e-4 of it (5e-9, which moves positions by ~1e-12 m, far too little to
change a decision).
"""
import ctypes as C
import sys

import numpy as np
import pytest

import scenes
from mps_amd import abi as A
from oracle.oracle import Oracle, make_config

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Problem:
    def __init__(self, lib, name, steps, seed=7, cdr=0.5, sigma=0.01):
        self.lib = lib
        self.steps = steps
        sc = scenes.CONFIGS[name]
        self.dim = sc["dim"]
        self.base = scenes.params_for(name, cg_tolerance=1e-13, cg_max_iter=20000, viscosity=1e-6,
                                      collision_distance_ratio=cdr)
        self.pos, vel, self.typ = scenes.dambreak(**sc)
        rng = np.random.default_rng(seed)
        fl = self.typ == 0
        self.u0 = np.zeros_like(vel)
        self.u0[fl] = rng.normal(0.0, sigma, size=(fl.sum(), self.dim))
        lib.set_option("tol_mode", 1)
        cfg = self.cfg()
        lib.set_config(cfg)
        s = A.HostScene(lib.abi, self.pos, self.u0, self.typ)
        lib.particle_system_calc_initial_params(s.ptr())
        self.n0, self.lam = s.ps.n0, s.ps.lambda_
        tpos, ttyp = self.final_positions(nu=5e-5)
        off = np.random.default_rng(seed + 1).normal(0.0, 1e-3, size=tpos.shape)
        self.target = (tpos + off * (ttyp == 0)[:, None], ttyp)

    def cfg(self, nu=None, g=None):
        p = dict(self.base)
        if nu is not None:
            p["viscosity"] = nu
        c = make_config(self.lib.abi, **p)
        if g is not None:
            for d in range(self.dim):
                c.gravity[d] = g[d]
        return c

    def scene(self, u0=None):
        s = A.HostScene(self.lib.abi, self.pos, self.u0 if u0 is None else u0, self.typ)
        s.ps.n0, s.ps.lambda_ = self.n0, self.lam
        return s

    def final_positions(self, nu=None, g=None, u0=None):
        cfg = self.cfg(nu, g)
        self.lib.set_config(cfg)
        s = self.scene(u0)
        self.lib.check(self.lib.mps_hip_run_steps(s.ptr(), self.steps), "run_steps")
        self.lib.check(self.lib.mps_hip_download(s.ptr(), None), "download")
        self.lib.mps_hip_release(s.ptr())
        return np.array(s.arr["pos"]), np.array(s.arr["type"])

    def loss(self, **kw):
        pos, typ = self.final_positions(**kw)
        fl = typ == 0
        tgt = self.target[0]
        return float(np.sum((pos[fl] - tgt[fl]) ** 2) / fl.sum())

    def oracle_loss(self, nu=None, g=None, u0=None):
        cfg = self.cfg(nu, g)
        sc = A.HostScene(A.types(self.dim), self.pos, self.u0 if u0 is None else u0, self.typ)
        sc.ps.n0, sc.ps.lambda_ = self.n0, self.lam
        sim = Oracle(self.dim).sim(tol_mode=1)
        sim.begin(sc, cfg)
        for _ in range(self.steps):
            sim.step(sc, cfg)
        pos, typ = np.array(sc.arr["pos"]), np.array(sc.arr["type"])
        fl = typ == 0
        return float(np.sum((pos[fl] - self.target[0][fl]) ** 2) / fl.sum())

    def adjoint(self, ckpt):
        cfg = self.cfg()
        self.lib.set_config(cfg)
        s = self.scene()
        n = s.num
        tgt = np.ascontiguousarray(self.target[0], dtype=np.float64)
        loss = np.zeros(1)
        gnu = np.zeros(1)
        gg = np.zeros(self.dim)
        gu = np.zeros((n, self.dim))
        self.lib.check(self.lib.mps_hip_adjoint(s.ptr(), self.steps, ckpt, _p(tgt), _p(loss), _p(gnu), _p(gg), _p(gu)),
                       "adjoint")
        return loss[0], gnu[0], gg, gu


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def _fd(loss, P, nu, g, hnu, hg, v, hu):
    out = {"nu": (loss(nu=nu + hnu) - loss(nu=nu - hnu)) / (2 * hnu)}
    for d in range(2):
        gp, gm = list(g), list(g)
        gp[d] += hg
        gm[d] -= hg
        out["g%d" % d] = (loss(g=gp) - loss(g=gm)) / (2 * hg)
    out["u"] = (loss(u0=P.u0 + hu * v) - loss(u0=P.u0 - hu * v)) / (2 * hu)
    return out


@pytest.mark.parametrize("name,steps,ckpt,cdr,rnu", [("dam2d_l0.01", 1, 1, 0.5, 1e-4),
                                                    ("dam2d_l0.01", 2, 1, 0.5, 1e-6),
                                                    ("dam2d_l0.01", 4, 2, 0.0, 1e-6),
                                                    ("dam2d_l0.01", 20, 6, 0.5, 1e-6),
                                                    ("dam2d_l0.005", 1, 1, 0.5, 1e-4)])
def test_adjoint_vs_finite_differences(lib2, name, steps, ckpt, cdr, rnu):
    P = Problem(lib2, name, steps, cdr=cdr)
    say("adjoint", steps, "steps, checkpoints every", ckpt, "collision ratio", cdr)
    loss, gnu, gg, gu = P.adjoint(ckpt)
    assert loss == pytest.approx(P.loss(), rel=1e-12)
    nu, g = P.base["viscosity"], [0.0, -9.81]
    rng = np.random.default_rng(11)
    v = np.zeros_like(P.u0)
    fl = P.typ == 0
    v[fl] = rng.normal(size=(fl.sum(), 2))
    steps_fd = (rnu * 5e-5, 1e-6 * 9.81, v, 1e-6 * np.linalg.norm(P.u0) / np.linalg.norm(v))
    ad = {"nu": gnu, "g0": gg[0], "g1": gg[1], "u": float(np.sum(gu * v))}
    fds = [("gpu", _fd(P.loss, P, nu, g, *steps_fd))]
    if steps >= 20:
        fds.append(("oracle", _fd(P.oracle_loss, P, nu, g, *steps_fd)))
    say("  loss %.12e" % loss)
    assert np.all(gu[~fl] == 0.0)
    for name, fd in fds:
        for key in ("nu", "g0", "g1", "u"):
            say("  %-4s adjoint %.10e  fd(%s) %.10e  rel %.2e" % (key, ad[key], name, fd[key], _rel(ad[key], fd[key])))
    for name, fd in fds:
        for key in ("nu", "g0", "g1", "u"):
            assert _rel(ad[key], fd[key]) < 1e-4, (name, key)


@pytest.mark.parametrize("name", ["dam2d_l0.01", "dam2d_l0.005"])
def test_adjoint_200_steps_vs_frozen_replay(lib2, name):
    """200 steps, checkpoints every 20.  Over this horizon a perturbation of
    1e-6 flips discrete decisions of the free forward (collisions, clamps,
    flags, list membership), so the differences are taken of the replay
    (mps_hip_replay_losses): the forward re-run with every decision of the
    recorded trajectory, the function the adjoint differentiates.  The
    replay's loss at the recorded parameters must equal the adjoint's.
    Central differences at relative step h = 1e-6 carry an h^2 error of a
    few 1e-4 at this horizon; Richardson's extrapolation from h and h/2
    cancels it.  dam2d_l0.005 is the differentiable workload itself
    (1.0 x 0.6 m tank, 0.3 x 0.6 m column at l0 = 0.005 m: 9,445
    particles, 7,200 fluid, dt = 5e-4 s)."""
    steps = 200
    P = Problem(lib2, name, steps)
    say("adjoint", steps, "steps, checkpoints every 20")
    loss, gnu, gg, gu = P.adjoint(20)
    nu, g = P.base["viscosity"], [0.0, -9.81]
    rng = np.random.default_rng(11)
    v = np.zeros_like(P.u0)
    fl = P.typ == 0
    v[fl] = rng.normal(size=(fl.sum(), 2))
    hnu, hg = 1e-6 * 5e-5, 1e-6 * 9.81
    hu = 1e-6 * np.linalg.norm(P.u0) / np.linalg.norm(v)
    sets = [(nu, g, P.u0)]
    for f in (1.0, 0.5):
        sets += [(nu + f * hnu, g, P.u0), (nu - f * hnu, g, P.u0),
                 (nu, [f * hg, -9.81], P.u0), (nu, [-f * hg, -9.81], P.u0),
                 (nu, [0.0, -9.81 + f * hg], P.u0), (nu, [0.0, -9.81 - f * hg], P.u0),
                 (nu, g, P.u0 + f * hu * v), (nu, g, P.u0 - f * hu * v)]
    nus = np.array([x[0] for x in sets])
    gs = np.ascontiguousarray(np.array([x[1] for x in sets], dtype=np.float64))
    us = np.ascontiguousarray(np.array([x[2] for x in sets], dtype=np.float64))
    losses = np.zeros(len(sets))
    P.lib.set_config(P.cfg())
    s = P.scene()
    tgt = np.ascontiguousarray(P.target[0], dtype=np.float64)
    P.lib.check(P.lib.mps_hip_replay_losses(s.ptr(), steps, _p(tgt), len(sets), _p(nus), _p(gs), _p(us),
                                            _p(losses)), "replay")
    say("  loss %.12e  replay %.12e" % (loss, losses[0]))
    assert losses[0] == pytest.approx(loss, rel=1e-12)
    ad = {"nu": gnu, "g0": gg[0], "g1": gg[1], "u": float(np.sum(gu * v))}
    hs = {"nu": hnu, "g0": hg, "g1": hg, "u": hu}
    errs = {}
    for k, key in enumerate(ad):
        d1 = (losses[1 + 2 * k] - losses[2 + 2 * k]) / (2 * hs[key])
        d2 = (losses[9 + 2 * k] - losses[10 + 2 * k]) / hs[key]
        rich = (4 * d2 - d1) / 3
        errs[key] = _rel(ad[key], rich)
        say("  %-3s adjoint %.10e  central %.10e (rel %.2e)  richardson %.10e (rel %.2e)"
            % (key, ad[key], d1, _rel(ad[key], d1), rich, errs[key]))
    for key, e in errs.items():
        assert e < 1e-4, key


def test_adjoint_bitwise_repeatable(lib2):
    """Two gradient runs give the same bits: every sum of the reverse sweep,
    the collision partners' shares included, runs in a fixed order.  The
    initial velocities (sigma 0.3 m/s) make collisions fire: without the
    collision rule (distance ratio 0) the gradient differs."""
    P = Problem(lib2, "dam2d_l0.01", 20, sigma=0.3)
    say("adjoint repeatability, 20 steps, sigma 0.3")
    a = P.adjoint(6)
    b = P.adjoint(6)
    assert a[0] == b[0]
    assert a[1] == b[1]
    assert np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64))
    assert np.array_equal(a[3].view(np.uint64), b[3].view(np.uint64))
    Q = Problem(lib2, "dam2d_l0.01", 20, sigma=0.3, cdr=0.0)
    c = Q.adjoint(6)
    say("  grad nu with collisions %.12e, without %.12e" % (a[1], c[1]))
    assert not np.array_equal(a[3], c[3])
