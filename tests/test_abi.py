"""The C-ABI library loads and exports every entry point include/frw.h
declares (no device calls)."""
import ctypes
import os
import re
import subprocess

import frw
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "frw.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(frw_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_python_export_list():
    assert _declared() == sorted(frw.EXPORTS)


def test_library_exports_all_declared_symbols():
    lib = frw.lib()
    for name in _declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", frw.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\b(frw_[a-z_0-9]+)\b", out))
    assert set(_declared()) <= exported


def test_version_and_error_string():
    lib = frw.lib()
    assert b"gfx950" in lib.frw_version()
    rc = lib.frw_layout_load(b"/nonexistent/layout", 0, ctypes.byref(ctypes.c_void_p()))
    assert rc == frw.E_IO
    assert b"cannot open" in lib.frw_last_error_string()


def test_shim_symbols_exported():
    """The reference-signature C++ entry points (include/randomwalk.h)."""
    out = subprocess.run(["nm", "-DC", "--defined-only", frw.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for sym in ["runRandomWalk(char*, int, int, bool, bool, int)",
                "getAvgPathLen(char*, int, int, bool, int)",
                "getAvgPathLenCUDA(char*, int, int, bool, int)",
                "randomWalkCudaWrapper(int, QuadTreeManager*, unsigned int*, RectHost*, unsigned long long)",
                "freeQuadTreeManager(QuadTreeManager*)", "countAvg(unsigned int*, int)",
                "Parser::parse(std::__cxx11::basic_string<char, std::char_traits<char>, std::allocator<char> > const&)",
                "initCuda(int, char**)",
                "createTree(Tree*, std::vector<RectHost, std::allocator<RectHost> > const&)",
                "RandomWalk(RectHost const&, Tree*, int&, double*, int)", "getIndex(double*, double)",
                "Tree::Tree(int, int, RectHost const&)", "Tree::insert(RectHost const&)",
                "Tree::checkCollisions(RectHost const&, RectHost const&)",
                "Tree::checkCollisons(point, RectHost&)", "Tree::drawBiggestSquareAtPoint(point)",
                "Tree::creatGaussianSurfFrom(RectHost const&, double)", "Tree::isInBounds(point const&)",
                "Tree::isInBounds(RectHost const&)", "Tree::clear()"]:
        assert sym in out, sym
