"""Host-side checks of the C++ case API (include/focalors/case_api.h): the
geometry and variable bookkeeping of base/domain/geometry2d.cpp and
base/domain/variable2d.cpp, run as a compiled program that never touches a
device."""
import case_programs


def test_case_api_geometry_host():
    out = case_programs.run("case_geometry")
    assert out.strip() == "ok", out
