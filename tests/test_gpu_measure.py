"""The measurement hooks with no reference counterpart (suli_ps_copy_bench,
suli_ps_read_bench): positive times, rates below the HBM peak, argument
checks."""
import pytest

pytestmark = pytest.mark.gpu

PEAK_GBS = 8000.0


@pytest.fixture(scope="module")
def solver():
    from suli_amd.solver import PressureSolver
    s = PressureSolver(verbose=False)
    yield s
    s.close()


def test_copy_and_read_rates(solver):
    nbytes = 512 << 20   # twice the 256 MB MALL: the rates are HBM rates
    copy_ms = solver.copy_bench(nbytes, reps=5)
    read_ms = solver.read_bench(nbytes, reps=5)
    print(f"copy {2 * nbytes / copy_ms / 1e6:.0f} GB/s, read {nbytes / read_ms / 1e6:.0f} GB/s", flush=True)
    assert copy_ms > 0 and read_ms > 0
    # neither moves its bytes faster than the HBM peak, and a read is no
    # slower than a copy of the same bytes
    assert 2 * nbytes / (copy_ms * 1e-3) / 1e9 < PEAK_GBS
    assert nbytes / (read_ms * 1e-3) / 1e9 < PEAK_GBS
    assert read_ms < copy_ms


def test_hooks_reject_bad_arguments(solver):
    from suli_amd.solver import SuliPsError
    with pytest.raises(SuliPsError):
        solver.read_bench(8, reps=1)
    with pytest.raises(SuliPsError):
        solver.copy_bench(1 << 20, reps=0)
