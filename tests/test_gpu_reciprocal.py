"""The tap projection's three-instruction reciprocal (recip_refined: one
Newton step from v_rcp_f32) equals the correctly rounded 1.0f / w for every
float w with |w| in [2^-40, 2^40), both signs, on this GPU: the exhaustive
check dpe_rcp_check (built with the library) must report no mismatch.  The
windows that use it are exactly those whose w provably stays in that range
(refined_division_ok)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "dpe-mvs-fast_amd", "dpe_rcp_check")


def test_recip_refined_is_correctly_rounded_exhaustively():
    assert os.path.exists(CHECK), "dpe_rcp_check is missing: build it with __graft_entry__.build()"
    r = subprocess.run([CHECK], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and " 0 mismatches" in r.stdout
