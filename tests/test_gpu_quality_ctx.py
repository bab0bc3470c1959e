"""The quality_ctx parity checks again in the GPU session, so the record of
the GPU test run holds them too. quality_ctx is a host coder (it uses no
device); these are the same functions as the CPU suite's."""
import pytest

import test_quality_ctx as T
import test_quality_ctx_pins as P

pytestmark = pytest.mark.gpu


def test_gpu_session_quality_ctx_hand_derived():
    T.test_hand_derived_one_symbol()
    T.test_hand_derived_two_symbols()
    P.test_hand_derived_context_switches()
    P.test_hand_derived_rescale()


def test_gpu_session_quality_ctx_reference_cases():
    T.test_range_coder_cases_of_the_reference()
    T.test_roundtrips_of_the_reference()
    T.test_error_cases_of_the_reference()
    T.test_library_matches_oracle_byte_for_byte()
