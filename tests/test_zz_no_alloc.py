"""The allocation-free trajectory entry point (hhg_trajectories_noalloc)
against hhg_trajectories, through the C ABI: bitwise equal outputs on two
calls in a row, the second with fewer points, and a request beyond the static
workspace rejected with INVALID_ARGUMENT."""
import numpy as np
import pytest

from tests.problems import CONFIG1, DIRAC_SMALL, HONEYCOMB_SMALL

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pr", [DIRAC_SMALL, HONEYCOMB_SMALL], ids=lambda p: p.name)
def test_trajectories_noalloc_matches_allocating_entry_point(hhg, pr):
    print(f"[case] trajectories_noalloc {pr.name}", flush=True)
    n_k = len(pr.o_grid()[0])
    with pr.h_plan(hhg) as plan:
        for sel in (np.arange(0, n_k, 3), np.array([n_k - 1, 5, 17])):
            ref = plan.trajectories(sel)
            got = plan.trajectories_noalloc(sel)
            assert got.shape == ref.shape == (len(sel), pr.time[2] + 1, 3)
            assert np.array_equal(got, ref)
            assert np.all(np.isfinite(got))


def test_trajectories_noalloc_rejects_requests_beyond_workspace(hhg):
    """4097 samples x 3 doubles per point: 171 points fit the 16 MiB
    workspace, 200 do not; the call fails before any launch and the next call
    still works."""
    print("[case] trajectories_noalloc size limit", flush=True)
    with CONFIG1.h_plan(hhg) as plan:
        with pytest.raises(hhg.HHGError) as ei:
            plan.trajectories_noalloc(np.arange(200))
        assert ei.value.status == hhg.ERROR_INVALID_ARGUMENT
        sel = np.arange(0, 8192, 64)[:100]
        assert np.array_equal(plan.trajectories_noalloc(sel), plan.trajectories(sel))
