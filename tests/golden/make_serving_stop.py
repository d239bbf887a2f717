"""Writes tests/golden/serving_rid19_stop.npz: a served request whose steady
run stops (a face temperature outside the thermo polynomial range, the
reference's `stop` in ComputeGasParameter) and what the compiled reference
and the C oracle do on it.

The request is rid 19 of the serving workload's seeded Poisson stream
(rocketcombustor_amd.serving.poisson_requests(24, 20.0), seed 4: a
73,960-cell nozzle/chamber case) in the block-multicolour sweep order.
The reference (oracle/_ref/rc_ref_dump, built by oracle/ref_build.sh) and
the oracle run up to NSTEPS steps from the same inputs; the fixture keeps
the cell permutation, every 10th step_residual row of both, and the step
(and, for the oracle, the face) where each stops, 0 when it does not.

Run:  python tests/golden/make_serving_stop.py   (about 15 minutes on one
core per code; needs the reference sources in this container)
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
import tempfile
import threading

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from rocketcombustor_amd import native, nozzle, serving  # noqa: E402

RID, NSTEPS, EVERY = 19, 5000, 10
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "rc_ref_dump")
REF_DECKS = "/root/reference/src/combustion"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "serving_rid19_stop.npz")


def run_reference(case, res):
    tmp = tempfile.mkdtemp(prefix="rcstop_")
    try:
        nozzle.write_reference_case(tmp, case, nstep_max=NSTEPS, extra_decks=REF_DECKS)
        os.makedirs(os.path.join(tmp, "out"))
        r = subprocess.run([REF_BIN, str(NSTEPS), "0", "out", "2"], cwd=tmp, capture_output=True, text=True)
        rows = np.fromfile(os.path.join(tmp, "out", "resid.bin")).reshape(-1, case.mech.nvar + 1)
        stopped = not os.path.exists(os.path.join(tmp, "out", "final_PA.bin"))
        res["ref_rows"] = rows
        res["ref_stop_step"] = len(rows) + 1 if stopped else 0
        res["ref_message"] = r.stdout.strip().splitlines()[-2:] if stopped else []
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main() -> None:
    from oracle import pyoracle
    req = serving.poisson_requests(24, 20.0)[RID]
    base = req.spec.build()
    perm, _ = native.order_multicolor(base.mesh.cells)
    case = base.permuted(perm)
    res = {}
    th = threading.Thread(target=run_reference, args=(case, res))
    th.start()
    o = pyoracle.OracleCase(case)
    rows, stop_step, stop_face = [], 0, 0
    for s in range(1, NSTEPS + 1):
        try:
            o.stage()
        except RuntimeError as e:
            stop_step = s
            m = re.search(r"face (\d+)", str(e))
            stop_face = int(m.group(1)) if m else -1
            break
        rows.append(o.residual())
    th.join()
    sp = req.spec
    np.savez_compressed(
        OUT, rid=RID, spec=np.array([sp.nx, sp.ny, sp.p0, sp.t0, sp.phi, sp.throat_ratio]),
        perm=perm.astype(np.int32), every=EVERY, nsteps=NSTEPS,
        ref_resid=res["ref_rows"][::EVERY], ref_stop_step=res["ref_stop_step"],
        oracle_resid=np.array(rows)[::EVERY], oracle_stop_step=stop_step, oracle_stop_face=stop_face)
    print("rid %d cells %d: reference stop %d, oracle stop %d (face %d); %.1f kB" % (
        RID, case.mesh.ne, res["ref_stop_step"], stop_step, stop_face, os.path.getsize(OUT) / 1024))


if __name__ == "__main__":
    main()
