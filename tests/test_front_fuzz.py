"""Malformed input never crashes the front end (CPU): mutated reference models and variants
files either parse or fail with a parse / limit error, and nesting deeper than the
expression-height bound (10,000 levels, Bison's default parser stack) is an error, not a
stack overflow."""
import random

import pytest

from tests.conftest import model_text
from upmabossg_amd import MbgError, Model

ALPHABET = "(){};=&|!^?:@$.,[]0123456789abcXYZ \n\t-+*/<>\"'#"


def _mutate(rng, text):
    t = list(text)
    for _ in range(rng.randint(1, 6)):
        if not t:
            break
        pos = rng.randrange(len(t))
        kind = rng.randrange(3)
        if kind == 0:
            del t[pos:pos + rng.randint(1, 12)]
        elif kind == 1:
            t[pos:pos] = t[pos:pos + rng.randint(1, 12)]
        else:
            t[pos] = rng.choice(ALPHABET)
    return "".join(t)


@pytest.mark.parametrize("name", ["metastasis", "cellcycle", "sizek"])
def test_mutated_models_parse_or_fail_cleanly(name):
    b0, c0 = model_text(name)
    rng = random.Random(name)
    parsed = 0
    for _ in range(300):
        b, c = (_mutate(rng, b0), c0) if rng.random() < 0.6 else (b0, _mutate(rng, c0))
        try:
            m = Model(b, c)
        except MbgError as e:
            assert e.status in (1, 2), e
            continue
        parsed += 1
        assert m.num_nodes > 0
        m.generate_source()
        m.transition_rates([0] * ((m.num_nodes + 63) // 64))
        text = f"variant v {{ {m.names()[0]}.knockout; {m.names()[-1]}.rate_up *= 0.5; }}"
        try:
            m.variants(_mutate(rng, text))
        except MbgError as e:
            assert e.status == 1
    assert parsed > 0


def test_nesting_beyond_the_bound_is_an_error():
    for e in ["(" * 20000 + "A" + ")" * 20000, "!" * 100000 + "A", " & ".join(["A"] * 50000)]:
        with pytest.raises(MbgError) as x:
            Model(f"node A {{ logic = {e}; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }}", "")
        assert "nested too deeply" in str(x.value)
    d = 9000
    m = Model(f"node A {{ logic = {'(' * d}!A{')' * d}; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }}", "")
    assert list(m.transition_rates([0])) == [1.0] and list(m.transition_rates([1])) == [1.0]
