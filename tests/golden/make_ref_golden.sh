#!/bin/sh
# Regenerates the reference-code golden vectors (ref_plan.json, ref_state.json,
# ref_forward_g12.json, ref_forward_g20.json) from the reference's own
# host-only C++ built by oracle/ref_build.sh.  Needs the reference sources
# (PGW_REFERENCE, default /root/reference); the fixtures it writes are data.
set -e
here=$(cd "$(dirname "$0")" && pwd)
root=$(cd "$here/../.." && pwd)
sh "$root/oracle/ref_build.sh"
bin="$root/oracle/_ref/ref_vectors"
"$bin" plan > "$here/ref_plan.json"
"$bin" state > "$here/ref_state.json"
python3 - "$bin" "$here" <<'PY'
import json, subprocess, sys
bin_, here = sys.argv[1], sys.argv[2]
for genus, sectors in ((12, range(6)), (20, (0, 1, 9))):
    runs = []
    for s in sectors:
        out = subprocess.run([bin_, "apply", f"{here}/g{genus}.json", str(s)], check=True,
                             capture_output=True, text=True).stdout
        rec = json.loads(out)
        rec["instance"] = f"g{genus}.json"
        runs.append(rec)
    with open(f"{here}/ref_forward_g{genus}.json", "w") as f:
        json.dump({"source": "cpu_eliminated_forward / cpu_apply_target_preconditioner of the reference "
                             "(cuda/cuprym_cyclic.cu:1685-1799) on deterministic_dense_block inputs",
                   "sectors": runs}, f, indent=1)
PY
