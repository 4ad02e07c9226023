"""The reference's own callers compile against the drop-in headers.

SfM/match_sequence/src/match_sequence.cpp and
SfM/match_using_voctree/src/match_using_voctree.cpp, with only the three cuBLAS
lines deleted (the integration guide's step 3), are syntax-checked by g++ with
include/ ahead of V3D/include, so include/CUDA/v3d_cudamatching.h (SIFT_Matching)
and include/Base/v3d_voctree.h (VocabularyTree, LeafDocuments) replace the
reference's headers and every use the drivers make of them type-checks.  The
sources are read from the reference tree at test time into a temporary
directory (nothing of them is kept here); the test is skipped where the
reference tree is absent.  -fpermissive and V3DLIB_ENABLE_SQLITE3 are the
reference's own build assumptions (its serialization header predates two-phase
name lookup; its storage header forward-declares the SQLite types).
"""
import os
import subprocess

import pytest

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVERS = ["match_sequence", "match_using_voctree"]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "SfM")), reason="reference tree not present")
@pytest.mark.parametrize("name", DRIVERS)
def test_reference_driver_compiles_against_dropin_headers(name, tmp_path):
    src = os.path.join(REF, "SfM", name, "src", name + ".cpp")
    with open(src, errors="replace") as f:
        lines = f.readlines()
    kept = [line for line in lines if "cublas" not in line]
    assert len(lines) - len(kept) == 3  # the include, cublasInit(), cublasShutdown()
    (tmp_path / (name + ".cpp")).write_text("".join(kept))
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-w", "-fpermissive", "-DV3DLIB_ENABLE_SQLITE3", "-H",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(REF, "V3D", "include"),
           "-I", os.path.join(REF, "V3D"), "-I", os.path.join(REF, "SfM", "reconstruction", "include"),
           name + ".cpp"]
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    included = [line.split()[-1] for line in r.stderr.splitlines() if line.startswith(".")]
    ours = os.path.join(ROOT, "include")
    assert os.path.join(ours, "CUDA", "v3d_cudamatching.h") in included
    if name == "match_using_voctree":
        assert os.path.join(ours, "Base", "v3d_voctree.h") in included
    assert not any(p.startswith(REF) and p.endswith(("v3d_cudamatching.h", "v3d_voctree.h")) for p in included)


SUBCLASS = r"""
#include <utility>
#include <vector>
#include "CUDA/v3d_cudamatching.h"
// a subclass reading the protected device members, and member pointers of the
// guided calls' exact types (v3d_cudamatching.h:48-66)
struct Sub : V3D_CUDA::SIFT_Matching
{
   float *scores() const { return _d_scores; }
   float *features() const { return _d_leftFeatures ? _d_leftFeatures : _d_rightFeatures; }
   float *positions() const { return _d_leftPositions ? _d_leftPositions : _d_rightPositions; }
   bool guided() const { return _useGuidedMatching; }
};
typedef void (V3D_CUDA::SIFT_Matching::*Guided)(V3D::Matrix3x3d const &, float, V3D::Vector2f const *,
                                                V3D::Vector2f const *, std::vector<std::pair<int, int> > &, float,
                                                float);
Guided g1 = &V3D_CUDA::SIFT_Matching::runGuidedMatching;
Guided g2 = &V3D_CUDA::SIFT_Matching::runGuidedHomographyMatching;
int main() { Sub s; return s.guided() ? 1 : 0; }
"""


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "V3D")), reason="reference tree not present")
@pytest.mark.parametrize("headers", ["dropin", "reference"])
def test_subclass_and_member_pointers_compile(headers, tmp_path):
    """The same code compiles against the reference's header (with its CUDA switch
    on) and against the drop-in one."""
    (tmp_path / "sub.cpp").write_text(SUBCLASS)
    inc = (["-I", os.path.join(ROOT, "include")] if headers == "dropin" else ["-DV3DLIB_ENABLE_CUDA"])
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-w", "-fpermissive", *inc, "-I", os.path.join(REF, "V3D", "include"),
           "sub.cpp"]
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
