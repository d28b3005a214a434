"""Host logic of the served form that needs no GPU: the precision modes a Server accepts,
request re-timing, and the request descriptors the native loop reads."""
import numpy as np
import pytest

from minifem.serve import Request, Server, retime


def test_server_rejects_f32coef_precision():
    """MF_PREC_MIXED_F32COEF is picked per request; a server in that mode would read
    float64 coefficient fields as float."""
    with pytest.raises(ValueError):
        Server([], 1024, precision=2)


def test_retime_keeps_payloads():
    reqs = [Request(r, 0.001 * r, r % 3, "coef", np.full(4, float(r))) for r in range(50)]
    for r in reqs:
        r.host = object()
    out = retime(reqs, 1e5, seed=7)
    assert [q.rid for q in out] == [r.rid for r in reqs]
    assert all(q.host is r.host and q.mesh == r.mesh for q, r in zip(out, reqs))
    t = np.array([q.t_arrival for q in out])
    assert np.all(np.diff(t) > 0) and t[-1] < 0.01  # 50 arrivals at 100,000/s


def test_request_descriptor_layout(tmp_path):
    """mf_request, mf_server_mesh, mf_server_stats and mf_plan_params in ctypes match
    include/minifem.h: sizes and field offsets printed by a C program built against the
    header."""
    import ctypes
    import os
    import shutil
    import subprocess
    from minifem._lib import PlanParams, Request as RequestDesc, ServerMesh, ServerStats
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    inc = os.path.join(os.path.dirname(__file__), "..", "include")
    src = tmp_path / "layout.c"
    src.write_text("""#include <stddef.h>
#include <stdio.h>
#include "minifem.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(mf_request), offsetof(mf_request, payload),
         offsetof(mf_request, payload_bytes), offsetof(mf_request, t_arrival),
         offsetof(mf_request, values_out), offsetof(mf_request, prec_out));
  printf("%zu %zu %zu %zu %zu\\n", sizeof(mf_server_mesh), offsetof(mf_server_mesh, nvalues),
         offsetof(mf_server_mesh, nprec), offsetof(mf_server_mesh, ncoord),
         offsetof(mf_server_mesh, ncoef));
  printf("%zu %zu\\n", sizeof(mf_server_stats), offsetof(mf_server_stats, makespan_s));
  printf("%zu %zu %zu\\n", sizeof(mf_plan_params), offsetof(mf_plan_params, id_chunks),
         offsetof(mf_plan_params, stage_precision));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.check_call([cc, "-I", inc, str(src), "-o", str(exe)])
    lines = [list(map(int, ln.split())) for ln in subprocess.check_output([str(exe)]).decode().split("\n") if ln]
    R = RequestDesc
    assert lines[0] == [ctypes.sizeof(R), R.payload.offset, R.payload_bytes.offset,
                        R.t_arrival.offset, R.values_out.offset, R.prec_out.offset]
    assert lines[1] == [ctypes.sizeof(ServerMesh), ServerMesh.nvalues.offset,
                        ServerMesh.nprec.offset, ServerMesh.ncoord.offset, ServerMesh.ncoef.offset]
    assert lines[2] == [ctypes.sizeof(ServerStats), ServerStats.makespan_s.offset]
    assert lines[3] == [ctypes.sizeof(PlanParams), PlanParams.id_chunks.offset,
                        PlanParams.stage_precision.offset]
