"""Models whose fast nodes read more than 20 nodes (logic programs instead of truth tables):
random expression trees over 21-40 inputs that use every operator a program supports
(& | ^ ! ?:, TRUE/FALSE, numbers, $constants, @aliases, unary minus, parentheses)."""
import numpy as np


def _tree(rng, pool, depth):
    r = rng.random()
    if depth <= 0 or r < 0.15:
        v = pool[int(rng.integers(len(pool)))]
        return ("!" if rng.random() < 0.3 else "") + v
    k = rng.random()
    if k < 0.1:
        return f"({_tree(rng, pool, depth - 1)} ? {_tree(rng, pool, depth - 1)} : {_tree(rng, pool, depth - 1)})"
    if k < 0.15:
        return f"!({_tree(rng, pool, depth - 1)})"
    if k < 0.18:
        return f"-({_tree(rng, pool, depth - 1)})"
    op = ["&", "|", "^"][int(rng.integers(3))]
    return f"({_tree(rng, pool, depth - 1)} {op} {_tree(rng, pool, depth - 1)})"


def wide_model(n=48, wide=12, seed=1, discrete=False, sample_count=2000, max_time=20.0, time_tick=0.5):
    """(bnd, cfg): n nodes; the first `wide` read 21-40 other nodes through random trees (half
    of them via an @alias attribute, some reading themselves); the rest are 2-3 input nodes."""
    rng = np.random.default_rng(seed)
    names = [f"w{i}" for i in range(n)]
    bnd = []
    for i in range(n):
        if i < wide:
            k = int(rng.integers(21, 41))
            ins = list(rng.choice(np.delete(np.arange(n), i), size=k, replace=False))
            if rng.random() < 0.3:
                ins.append(i)
            pool = [names[j] for j in ins] + ["TRUE", "FALSE", "$c1", "$c0", "0.5"]
            # every input appears at least once: AND/OR/XOR of the required literals with a tree
            parts = [names[j] if rng.random() < 0.5 else "!" + names[j] for j in ins]
            chain = parts[0]
            for p in parts[1:]:
                chain = f"({chain} {['&', '|', '^'][int(rng.integers(3))]} {p})"
            expr = f"({chain}) {['|', '^', '&'][int(rng.integers(3))]} ({_tree(rng, pool, 5)})"
            if rng.random() < 0.5:
                bnd.append(f"node {names[i]} {{\n  inner = {expr};\n  logic = @inner;\n"
                           f"  rate_up = @logic ? $u{i} : 0;\n  rate_down = @logic ? 0 : $d{i};\n}}\n")
                continue
        else:
            k = int(rng.integers(2, 4))
            ins = list(rng.choice(np.delete(np.arange(n), i), size=k, replace=False))
            expr = " & ".join(("!" if rng.random() < 0.5 else "") + names[j] for j in ins)
        bnd.append(f"node {names[i]} {{\n  logic = {expr};\n  rate_up = @logic ? $u{i} : 0;\n"
                   f"  rate_down = @logic ? 0 : $d{i};\n}}\n")
    cfg = ["$c1 = 1;", "$c0 = 0;"]
    up = np.exp(rng.uniform(np.log(0.2), np.log(5.0), n)).astype(np.float32)
    down = np.exp(rng.uniform(np.log(0.2), np.log(5.0), n)).astype(np.float32)
    for i in range(n):
        cfg += [f"$u{i} = {float(up[i]):.9g};", f"$d{i} = {float(down[i]):.9g};",
                f"{names[i]}.istate = {float(rng.uniform()):.6f};"]
        if i >= 6:
            cfg.append(f"{names[i]}.is_internal = 1;")
    cfg += [f"sample_count = {sample_count};", f"max_time = {max_time:g};", f"time_tick = {time_tick:g};",
            f"discrete_time = {int(discrete)};", "seed_pseudorandom = 3;"]
    return "".join(bnd), "\n".join(cfg) + "\n"
