"""CPU check that the C++ gls:: API test program links against libglsx.so."""
import os


def test_cpp_program_built():
    import glscv_amd  # noqa: F401

    here = os.path.dirname(os.path.abspath(__file__))
    assert os.path.exists(os.path.join(here, "cpp", "test_gls_api"))
