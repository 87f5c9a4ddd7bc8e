"""fl_generate_f64_noalloc (every device buffer in the library's static
device scratch, no allocation) against fl_generate_f64 on the same terrains,
bit for bit, through the C ABI: two calls in a row, the second with a smaller
terrain (the scratch keeps the first call's contents); a terrain too large
for the scratch is refused before anything runs."""
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_matches_generate():
    import noalloc_cases
    from fastlem_amd import _native
    from fastlem_amd.generator import context

    noalloc_cases.check_noalloc(_native.lib(), context())
