"""The oracle reproduces the committed golden fixtures bit for bit (digests) and
to float64 round-off (projections): a regression pin on the oracle itself."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_oracle_matches_golden(orc):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_golden
    ref = np.load(os.path.join(HERE, "golden", "siddon_cases.npz"))
    new = make_golden.build_cases()
    assert set(ref.files) == set(new)
    for k in ref.files:
        if k.endswith("_hash") or k.endswith("_count") or k in ("lors",) or k.endswith("_image") \
                or k.endswith("_values"):
            assert np.array_equal(ref[k], new[k]), k
        else:
            np.testing.assert_allclose(new[k], ref[k], rtol=1e-6, atol=1e-9, err_msg=k)


def test_golden_covers_every_branch(orc):
    # every branch id 0..10 of the marcher occurs in the fixtures
    import oracle
    ref = np.load(os.path.join(HERE, "golden", "siddon_cases.npz"))
    lors = ref["lors"]
    seen = set()
    shape, center, size = (32, 28, 16), [0.0, 0.0, 0.0], [64.0, 56.0, 32.0]
    for l in lors[:400]:
        for var in (0, 1):
            seen |= set(oracle.terms(l, shape, center, size, tof=(3.66, 649.2), variant=var)[1].tolist())
    assert seen == set(range(11))
