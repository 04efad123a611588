"""The oracle against closed-form golden values: energies and densities of
perfect and homogeneously strained BCC lattices under the Finnis-Sinclair Fe
potential, computed by direct analytic shell sums (tests/golden/
make_fs_lattice.py, no spline and no library code).  The tabulated functions
are exact cubics at every neighbour distance of these lattices, so the spline
reproduces them to rounding."""
import numpy as np
import pytest

import oracle as O
import workload as W

CASES = W.golden_fs_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_closed_form_lattice(case):
    c = W.strained_lattice(case["lattice_const"], case["strain"])
    st, T = c["st"], O.Tables(c["tab"], c["masses"])
    e_emb = O.rho_df(c["dom"], c["even"], c["odd"], T, W.RC, st, want_e=True)
    O.ghost_df(c["dom"], st)
    e_pair = O.force(c["dom"], c["even"], c["odd"], T, W.RC, st, want_e=True)
    li = st.local_index()
    n = li.size
    assert np.allclose(st.rho[li], case["rho"], rtol=1e-12, atol=0)
    assert abs(e_emb / n - case["embed"]) < 1e-11 * abs(case["embed"])
    assert abs(e_pair / n - case["pair_half"]) < 1e-11 * abs(case["embed"])
    assert abs((e_emb + e_pair) / n - case["energy"]) < 1e-11 * abs(case["energy"])
    assert np.abs(np.stack([st.fx[li], st.fy[li], st.fz[li]])).max() < 1e-11


def test_golden_energy_minimum_and_published_fit():
    e = {c["name"]: c["energy"] for c in CASES}
    # the published fit: E_coh = 4.28 eV at a = 2.8665 A, the minimum of the sampled constants
    assert abs(e["a2.8665"] + 4.28) < 1e-3
    assert e["a2.8665"] < min(e["a2.810"], e["a2.855"], e["a2.900"])
    # a volume-conserving strain of the equilibrium-near lattice costs energy
    assert e["a2.855_tetragonal"] > e["a2.855"]
