"""Seeded synthetic random Boolean networks, written as .bnd/.cfg text.

The same bytes feed the MaBoSSG CLI, the library and the CPU oracle.

Formula family (fixed): a node reads k distinct other nodes (k drawn per
configuration); each literal is x_j or !x_j with probability 1/2; literals are
combined left-deep, ((l1 op l2) op l3) ..., with op uniform in {&, |}.  Rates
use the reference's fast form, ``rate_up = @logic ? $u_X : 0`` and
``rate_down = @logic ? 0 : $d_X`` (src/generator.cpp:57-107), with the FP32
values printed to 9 significant digits so they parse back exactly.

Configurations (sizes of the workload each stands for):
  small       16 nodes, in-degree 2-3, rates 1, istate 0.5, all 16 readouts,
              10,000 trajectories, max_time 100, time_tick 1
  large       1,000 nodes, in-degree 2-4, rates log-uniform [0.1, 10], istate
              U[0, 1], readouts nodes 0-15, 1,048,576 trajectories,
              max_time 50, time_tick 0.5
  fixed_heavy 5,000 nodes, in-degree 1-3, 30% special nodes (half
              self-sustaining x = x | f, half absorbing rate_down = 0),
              262,144 trajectories, max_time 200, time_tick 2
  sweep       variants of 50-2,000 nodes in the style of `large`
  surrogate   200 nodes in the style of `large`
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Network:
    n: int
    logic: list            # logic expression text per node
    up: np.ndarray         # float32
    down: np.ndarray       # float32
    istate: np.ndarray     # float32
    readouts: list         # node indices that are non-internal
    max_time: float = 100.0
    time_tick: float = 1.0
    sample_count: int = 10000
    seed: int = 1
    discrete: bool = False
    names: list = field(default_factory=list)

    def __post_init__(self):
        if not self.names:
            self.names = [f"n{i}" for i in range(self.n)]

    def bnd(self) -> str:
        out = []
        for i, nm in enumerate(self.names):
            out.append(f"node {nm} {{\n  logic = {self.logic[i]};\n  rate_up = @logic ? $u_{nm} : 0;\n"
                       f"  rate_down = @logic ? 0 : $d_{nm};\n}}\n")
        return "".join(out)

    def cfg(self) -> str:
        ro = set(self.readouts)
        lines = []
        for i, nm in enumerate(self.names):
            lines.append(f"$u_{nm} = {_f32(self.up[i])};")
            lines.append(f"$d_{nm} = {_f32(self.down[i])};")
        for i, nm in enumerate(self.names):
            lines.append(f"{nm}.istate = {_f32(self.istate[i])};")
            if i not in ro:
                lines.append(f"{nm}.is_internal = 1;")
        lines += [f"sample_count = {self.sample_count};", f"max_time = {_f32(self.max_time)};",
                  f"time_tick = {_f32(self.time_tick)};", f"discrete_time = {int(self.discrete)};",
                  f"seed_pseudorandom = {self.seed};"]
        return "\n".join(lines) + "\n"

    def write(self, prefix: str):
        with open(prefix + ".bnd", "w") as f:
            f.write(self.bnd())
        with open(prefix + ".cfg", "w") as f:
            f.write(self.cfg())
        return prefix + ".bnd", prefix + ".cfg"

    def inputs(self):
        """Node indices each logic reads (for tests)."""
        import re
        idx = {nm: i for i, nm in enumerate(self.names)}
        return [sorted({idx[t] for t in re.findall(r"[A-Za-z_][A-Za-z_0-9]*", e) if t in idx}) for e in self.logic]


def _f32(v) -> str:
    # 9 significant digits identify an FP32 value exactly through strtod -> (float)
    return "%.9g" % float(np.float32(v))


def _formula(rng, names, i, k_lo, k_hi, n):
    k = int(rng.integers(k_lo, k_hi + 1))
    k = min(k, n - 1)
    if k <= 0:
        return names[i]
    choices = rng.choice(np.delete(np.arange(n), i), size=k, replace=False)
    lits = [("!" if rng.random() < 0.5 else "") + names[j] for j in choices]
    expr = lits[0]
    for lit in lits[1:]:
        op = "&" if rng.random() < 0.5 else "|"
        expr = f"({expr} {op} {lit})"
    return expr


def random_network(n, k_lo, k_hi, seed, rates="unit", istate="half", readouts=16, special_frac=0.0,
                   max_time=100.0, time_tick=1.0, sample_count=10000, sim_seed=1) -> Network:
    rng = np.random.default_rng(seed)
    names = [f"n{i}" for i in range(n)]
    logic = [_formula(rng, names, i, k_lo, k_hi, n) for i in range(n)]
    if rates == "unit":
        up = np.ones(n, np.float32)
        down = np.ones(n, np.float32)
    else:
        up = np.exp(rng.uniform(np.log(0.1), np.log(10.0), n)).astype(np.float32)
        down = np.exp(rng.uniform(np.log(0.1), np.log(10.0), n)).astype(np.float32)
    ist = np.full(n, 0.5, np.float32) if istate == "half" else rng.uniform(0, 1, n).astype(np.float32)
    if special_frac > 0:
        special = rng.choice(n, size=int(round(special_frac * n)), replace=False)
        half = len(special) // 2
        for i in special[:half]:       # self-sustaining: x = x | f
            logic[i] = f"{names[i]} | {logic[i]}"
        for i in special[half:]:       # absorbing: once ON, stays ON
            down[i] = 0.0
    ro = list(range(min(readouts, n)))
    return Network(n, logic, up, down, ist, ro, max_time, time_tick, sample_count, sim_seed, False, names)


def config_small(seed=1) -> Network:
    return random_network(16, 2, 3, seed, rates="unit", istate="half", readouts=16, max_time=100.0,
                          time_tick=1.0, sample_count=10000)


def config_large(seed=1, n=1000) -> Network:
    return random_network(n, 2, 4, seed, rates="loguniform", istate="uniform", readouts=16, max_time=50.0,
                          time_tick=0.5, sample_count=1048576)


def config_fixed_heavy(seed=1, n=5000) -> Network:
    return random_network(n, 1, 3, seed, rates="loguniform", istate="uniform", readouts=16, special_frac=0.3,
                          max_time=200.0, time_tick=2.0, sample_count=262144)


def config_surrogate(seed=1) -> Network:
    return random_network(200, 2, 4, seed, rates="loguniform", istate="uniform", readouts=16, max_time=50.0,
                          time_tick=0.5, sample_count=4096)


def sweep_variants(count=256, seed=1, n_lo=50, n_hi=2000, traj_lo=1000, traj_hi=64000):
    """Variant networks: size, rate scale log-normal(0, 0.5), 0-3 knockouts, trajectory count."""
    rng = np.random.default_rng(seed)
    out = []
    for v in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        net = random_network(n, 2, 4, seed * 100003 + v, rates="loguniform", istate="uniform", readouts=16,
                             max_time=50.0, time_tick=0.5, sample_count=int(rng.integers(traj_lo, traj_hi + 1)))
        scale = np.float32(np.exp(rng.normal(0.0, 0.5)))
        net.up = (net.up * scale).astype(np.float32)
        net.down = (net.down * scale).astype(np.float32)
        for i in rng.choice(n, size=int(rng.integers(0, 4)), replace=False):
            net.up[i] = 0.0          # knockout: forced OFF
            net.istate[i] = 0.0
        out.append(net)
    return out


def ring_network(nodes=100, noninternals=5, signal_length=0, steps_to_fixed_point=0, sample_count=1000000,
                 time_tick=5.0, max_time=100.0, discrete=False, contiguous_noninternals=False):
    """The ring models of the reference's benchmark generator (data/generate-synth.py:5-100),
    as (bnd text, cfg text): a signal of s = signal_length (default int(sqrt(nodes))) ON nodes
    travels around a ring.  Node i turns ON when it is OFF, its s predecessors are ON and the
    node before them is OFF, and turns OFF when it and its s successors are ON, so its logic
    reads 2s + 1 other nodes (22 at 100 nodes, 64 at 1,000).  With steps_to_fixed_point > 0 a
    last node NFP (rate_up 1/steps, rate_down 0) switches every ring node ON for good, which
    makes all-ON a fixed point.  Unit rates; the first s nodes start ON; every
    (nodes // noninternals)-th node (or the first noninternals nodes when contiguous) and NFP
    are non-internal."""
    fixed = steps_to_fixed_point != 0
    ring = nodes - 1 if fixed else nodes
    s = signal_length if signal_length > 0 else int(np.sqrt(nodes))
    stride = 1 if contiguous_noninternals else nodes // noninternals

    def name(k):
        return f"N{k % ring}"

    bnd = []
    for i in range(ring):
        on_terms = [f"!N{i}"] + (["!NFP"] if fixed else []) + [name(i - j + ring) for j in range(1, s + 1)]
        on_terms.append("!" + name(i - s - 1 + ring))
        stay = " & ".join([f"N{i}"] + [name(i + j) for j in range(1, s + 1)])
        logic = f"({' & '.join(on_terms)}) | (!({stay}) & N{i}" + (" & !NFP" if fixed else "") + ")"
        if fixed:
            logic += " | (NFP)"
        bnd.append(f"Node N{i} {{\n    logic = {logic};\n    rate_up = @logic ? $u_N{i} : 0;\n"
                   f"    rate_down = @logic ? 0 : $d_N{i};\n}}\n")
    if fixed:
        bnd.append("Node NFP {\n    logic = !NFP;\n    rate_up = @logic ? $u_NFP : 0;\n    rate_down = @logic ? 0 : $d_NFP;\n}\n")
    cfg = [f"sample_count = {sample_count};", f"time_tick = {time_tick:g};", f"max_time = {max_time:g};",
           f"discrete_time = {int(discrete)};"]
    if fixed:
        cfg += [f"$u_NFP = {1.0 / steps_to_fixed_point!r};", "$d_NFP = 0;", "NFP.istate = 0;", "NFP.is_internal = 0;"]
    left = noninternals
    for i in range(ring):
        cfg += [f"$u_N{i} = 1;", f"$d_N{i} = 1;", f"N{i}.istate = {1 if i < s else 0};"]
        if i % stride == 0 and left > 0:
            cfg.append(f"N{i}.is_internal = 0;")
            left -= 1
        else:
            cfg.append(f"N{i}.is_internal = 1;")
    return "\n".join(bnd), "\n".join(cfg) + "\n"
