"""Pick, per three-centre class and precision, how the kernels stage pair-matrix entries ("lds" or
"dpp"), from autotuner logs of two builds that differ only in the staging mode.

Usage: python -m codegen.choose_stage LDS_LOG[,LDS_LOG...] DPP_LOG[,DPP_LOG...] > codegen/stage_choice.json
(run from the package directory).  Each log holds lines "autotune <la><lb><lc>_<prec> variant <v>: <t> ms"
as printed by engine.autotune; a class's time is its best variant, summed over the logs (workloads) that
time it.  "dpp" is chosen when it is at least 2 % faster; otherwise the class keeps "lds".
"""

import json
import re
import sys

_LINE = re.compile(r"autotune (\d{3}_f(?:32|64)) variant (\d+): ([\d.]+) ms")


def best_times(paths):
    total = {}
    for path in paths:
        best = {}
        with open(path) as f:
            for line in f:
                m = _LINE.search(line)
                if m:
                    k, t = m.group(1), float(m.group(3))
                    best[k] = min(best.get(k, t), t)
        for k, t in best.items():
            total[k] = total.get(k, 0.0) + t
    return total


def choose(lds_logs, dpp_logs, margin=0.98):
    from .gen_kernels import DPP_MAX_ENTRIES, pair_nnz
    lds, dpp = best_times(lds_logs), best_times(dpp_logs)

    def pick(k):
        la, lb = int(k[0]), int(k[1])
        if pair_nnz(la, lb, True) > DPP_MAX_ENTRIES:       # such rows always stay in LDS
            return "lds"
        return "dpp" if k in dpp and dpp[k] < margin * lds[k] else "lds"
    return {k: pick(k) for k in sorted(lds)}


if __name__ == "__main__":
    table = choose(sys.argv[1].split(","), sys.argv[2].split(","))
    print(json.dumps(table, indent=1, sort_keys=True))
