"""The build/smoke entry point's smoke(): one small cfmm2d_ and one small
hfmm2d_ solve on the device, each checked against the oracle."""
import pytest

pytestmark = pytest.mark.gpu


def test_smoke_entry_point():
    import __graft_entry__
    __graft_entry__.smoke()
