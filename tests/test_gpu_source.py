"""The reference's source known-answer tests (test/source.jl:5-232: the exact
distweights vectors, the PointSrc invariants and the PlaneSrc testset) run
again inside the GPU suite, so that the libmfdfd legs of the point-dipole and
plane-source rows also pass in the run on the MI355X.  The same functions are
collected, unmarked, by test_source.py in the CPU suite."""
import pytest

from test_source import (test_distweights_kat, test_distweights_n1, test_planesrc_capi_matches_oracle,  # noqa: F401
                         test_planesrc_reference_testset, test_planesrc_rejects, test_pointsrc_capi_matches_oracle,
                         test_pointsrc_invariants)

pytestmark = pytest.mark.gpu
