"""Pick, per three-centre class and precision, how the kernels stage pair-matrix entries, from autotuner
logs of builds that differ only in the staging mode:

* "lds": the workgroup's pair-matrix rows in LDS, each entry read by a broadcast LDS load where used;
* "dpp": the rows in registers, 64 entries per lane-parallel load, each entry broadcast by a DPP move
  (only for rows of at most DPP_MAX_ENTRIES entries);
* "ldsdpp": the rows in LDS, read 16 entries per lane-parallel LDS load into rotating registers two
  groups ahead of use, each entry broadcast by a DPP move (any row length).

Usage (from the package directory):
  python -m codegen.choose_stage LDS_LOG[,...] DPP_LOG[,...] > codegen/stage_choice.json
  python -m codegen.choose_stage --ldsdpp BASE_LOG[,...] LDSDPP_LOG[,...] < old.json > codegen/stage_choice.json
The first form picks "lds" or "dpp"; the second moves the classes of an existing table that are on "lds"
to "ldsdpp", from the logs of a build of that table and of a build whose "lds" classes use "ldsdpp".  Each log holds lines "autotune <la><lb><lc>_<prec> variant <v>: <t> ms"
as printed by engine.autotune; a class's time is its best variant, summed over the logs (workloads) that
time it.  A mode other
than "lds" is chosen when it is at least 2 % faster than "lds"; otherwise the class keeps "lds".
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


def upgrade_lds(table, base_logs, ldsdpp_logs, margin=0.98):
    base, alt = best_times(base_logs), best_times(ldsdpp_logs)
    return {k: ("ldsdpp" if m == "lds" and k in base and k in alt and alt[k] < margin * base[k] else m)
            for k, m in table.items()}


if __name__ == "__main__":
    if sys.argv[1] == "--ldsdpp":
        table = upgrade_lds(json.load(sys.stdin), sys.argv[2].split(","), sys.argv[3].split(","))
    else:
        table = choose(sys.argv[1].split(","), sys.argv[2].split(","))
    print(json.dumps(table, indent=1, sort_keys=True))
