"""om_field_statistics_noalloc through the C ABI: the allocation-free variant of
om_field_statistics (its reduction works in the library's static device array) must give exactly
the statistics the allocating entry point gives — maximum and minimum are exact operations and both
reductions have the same fixed shape — on two calls in a row, the second with a smaller field,
with a planted NaN, and on the parent and the interior of a field."""
import numpy as np
import pytest

from cases import convection, product_for

pytestmark = pytest.mark.gpu


def _both(model, field, with_halos):
    a = model.field_statistics(field.id, with_halos)
    b = model.field_statistics_noalloc(field.id, with_halos)
    print(f"field {field.id} halos={with_halos}: allocating {a} noalloc {b}", flush=True)
    return a, b


def test_noalloc_statistics_equal_the_allocating_entry_point():
    big = product_for(*convection(N=(48, 40, 24), float_type=np.float32, noise_velocity=1e-2, seed=4))
    small = product_for(*convection(N=(16, 16, 8), float_type=np.float64, noise_velocity=1e-2, seed=5))
    big.time_step(1.0, 2)
    small.time_step(1.0, 1)
    # the first call reduces a 54×46×31 parent, the second a 16×16×8 interior
    for model, field, halos in ((big, big.velocities.w, True), (small, small.tracers.b, False),
                                (big, big.tracers.b, False), (small, small.velocities.u, True)):
        a, b = _both(model, field, halos)
        assert a == b
        assert b[2] == 0
        ref = field.parent() if halos else field.interior()
        assert (b[0], b[1]) == (ref.max(), ref.min())
    # a planted NaN in a halo point is counted over the parent, not the interior
    p = small.velocities.v.parent()
    p[0, 1, 2] = np.nan
    p[3, 4, 5] = np.nan
    small.velocities.v.set_parent(p)
    a, b = _both(small, small.velocities.v, True)
    assert a == b and b[2] == 2
    a, b = _both(small, small.velocities.v, False)
    assert a == b and b[2] == 1


def test_noalloc_statistics_reject_an_unknown_field():
    small = product_for(*convection(N=(16, 16, 8), float_type=np.float32, seed=6))
    import oceananigans_mi355x as om
    with pytest.raises(om.OmError, match="unknown field"):
        small.field_statistics_noalloc(999, True)
