"""Write the committed golden fixtures (run in a container that has /root/reference; never at test time).

1. reference_cases.json + reference_cases.npz -- the reference's own integral test cases
   (reference molecular_integral/tests/test_{overlap,nuclear_attraction,two_center,three_center}_integral.py):
   inputs (geometry text, unit, spherical flag, omega, basis-set text, point charges, tolerance) read out
   of the test files as literal data, and the expected arrays from the data files beside them (PySCF
   output).  Named basis sets are stored as the element blocks the cases use.
2. ref_class_vectors.npz -- per-class outputs of the reference's C++ compiled from its own sources
   (oracle/_ref, see oracle/Makefile) on small seeded inputs: every 3c class (orbital l <= 6, aux l <= 9),
   2c class (l <= 9) and S/V class (l <= 6), spherical, omega 0 and 0.2.  These pin the classes whose
   PySCF fixtures the reference does not ship (all-angular three-centre data is missing upstream).

Usage: python tests/golden/make_fixtures.py
"""

import ast
import ctypes
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REF = pathlib.Path("/root/reference/electron_integral_playground")
TESTS = REF / "molecular_integral" / "tests"
BASIS_DIR = REF / "data" / "basis_sets"

KINDS = {"overlap": "S", "nuclear_attraction": "V", "two_center": "J2c", "three_center": "J3c"}


def _literal(node, env):
    if isinstance(node, ast.Name):
        return env[node.id]
    if isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name) and node.value.id == "LengthUnits":
        return node.attr
    return ast.literal_eval(node)


def element_blocks(text, elements):
    """Keep only the ``****``-terminated blocks of the listed elements."""
    out, keep, block = [], False, []
    for line in text.split("\n"):
        f = line.split()
        if f and f[0].startswith("!"):
            continue
        if not block and len(f) == 2 and f[1] == "0":
            keep = f[0] in elements
        if f or block:
            block.append(line)
        if f and f[0].startswith("****"):
            if keep:
                out.extend(block)
            block, keep = [], False
    return "\n".join(out) + "\n"


def extract_cases():
    cases = []
    arrays = {}
    for kind_file, key in KINDS.items():
        path = TESTS / f"test_{kind_file}_integral.py"
        tree = ast.parse(path.read_text())
        consts = {}
        for node in tree.body:
            if isinstance(node, ast.Assign) and isinstance(node.value, ast.Constant):
                consts[node.targets[0].id] = node.value.value
        permissive = "permissive_xyz_reader" in path.read_text()
        for fn in [n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name.startswith("test_")]:
            case = {"name": fn.name, "kind": key, "permissive": permissive, "spherical": True, "omega": 0.0,
                    "unit": "ANGSTROM", "charges": None}
            env = dict(consts)
            for node in ast.walk(fn):
                if isinstance(node, ast.Assign):
                    tgt = node.targets[0]
                    val = node.value
                    if isinstance(tgt, ast.Attribute) and tgt.attr == "spherical_basis":
                        case["spherical"] = bool(ast.literal_eval(val))
                    elif isinstance(tgt, ast.Name) and tgt.id == "omega":
                        case["omega"] = float(ast.literal_eval(val))
                    elif isinstance(tgt, ast.Name) and isinstance(val, ast.Call):
                        fname = getattr(val.func, "id", getattr(val.func, "attr", ""))
                        if fname == "read_xyz_content":
                            case["xyz"] = _literal(val.args[0], env)
                            for kw in val.keywords:
                                if kw.arg == "unit":
                                    case["unit"] = _literal(kw.value, env)
                        elif fname in ("read_basis_content", "read_basis"):
                            which = "aux_basis" if "auxiliary" in tgt.id else "basis"
                            if kind_file == "two_center":
                                which = "aux_basis"
                            arg = _literal(val.args[0], env)
                            case[which] = {"text": arg} if fname == "read_basis_content" else {"name": arg}
                        elif fname == "array" and tgt.id == "charge_position":
                            case["charges"] = _literal(val.args[0], env)
                        elif fname in ("loadtxt", "load"):
                            arg = val.args[0]
                            if isinstance(arg, ast.BinOp):
                                case["data_file"] = _literal(arg.right, env)
                        elif fname == "" and isinstance(val.func, ast.Attribute):
                            pass
                    elif isinstance(tgt, ast.Name) and tgt.id == "nuclear_position":
                        case["charges"] = "nuclei"
                    elif isinstance(tgt, ast.Name) and tgt.id in ("ref_J3c",) and isinstance(val, ast.Subscript):
                        call = val.value
                        case["data_file"] = _literal(call.args[0].right, env)
                elif isinstance(node, ast.Call) and getattr(node.func, "attr", "") == "assert_allclose":
                    for kw in node.keywords:
                        if kw.arg == "atol":
                            case["atol"] = float(_literal(kw.value, env))
            data_path = TESTS / f"reference_{kind_file}" / case.get("data_file", "missing")
            if not data_path.exists():
                print("skip (data not shipped upstream):", fn.name)
                continue
            if data_path.suffix == ".npz":
                with np.load(data_path, allow_pickle=False) as z:
                    arr = z["J3c"]
            else:
                arr = np.loadtxt(data_path)
            elements = set(l.split()[0] for l in case["xyz"].split("\n") if len(l.split()) == 4)
            for which in ("basis", "aux_basis"):
                if which in case and "name" in case[which]:
                    text = (BASIS_DIR / (case[which]["name"].lower() + ".gaussian.basis")).read_text()
                    case[which] = {"text": element_blocks(text, elements), "name": case[which]["name"]}
            arrays[fn.name] = arr
            cases.append(case)
    return cases, arrays


def class_vectors():
    """Per-class reference outputs from oracle/_ref on seeded inputs (spherical, two omegas)."""
    sys.path.insert(0, str(ROOT))
    from tests.helpers import REF_SO, dp, ip
    lib = ctypes.CDLL(str(REF_SO))
    rng = np.random.default_rng(20261203)
    out = {}

    def pair_block(la, lb):
        A = rng.uniform(-1, 1, 3); B = A + rng.uniform(-1.2, 1.2, 3)
        a = rng.uniform(0.4, 2.0, 2); b = rng.uniform(0.4, 2.0, 2)
        rows = []
        for x in a:
            for y in b:
                p = x + y
                rows.append((np.r_[(x * A + y * B) / p, p], np.r_[A, x], np.r_[B, y], rng.normal()))
        P, AA, BB, c = (np.ascontiguousarray(np.array([r[k] for r in rows])) for k in range(4))
        ni, nj = 2 * la + 1, 2 * lb + 1
        return P, AA, BB, c, np.zeros(4, np.int32), np.full(4, ni, np.int32), ni + nj

    for omega in (0.0, 0.2):
        for lb in range(7):
            for la in range(lb + 1):
                P, AA, BB, c, si, sj, nao = pair_block(la, lb)
                for lc in range(10):
                    C = np.ascontiguousarray(np.c_[rng.uniform(-1, 1, (1, 3)), rng.uniform(0.3, 2.5, 1)])
                    ck = np.array([rng.normal()]); ks = np.zeros(1, np.int32)
                    nk = 2 * lc + 1
                    J = np.zeros((nk, nao, nao))
                    rc = lib.three_center(la, lb, lc, dp(P), dp(AA), dp(BB), dp(c), ip(si), ip(sj), 4, dp(C), dp(ck),
                                          ip(ks), 1, dp(J), nao, nk, ctypes.c_bool(True), ctypes.c_double(omega))
                    assert rc == 0
                    tag = f"3c_{la}{lb}{lc}_w{omega}"
                    out[tag + "_J"] = J[:, :2 * la + 1, 2 * la + 1:]
                    out[tag + "_in"] = np.concatenate([P.ravel(), AA.ravel(), BB.ravel(), c, C.ravel(), ck])
                # S and V for this pair
                S = np.zeros((nao, nao))
                assert lib.overlap(la, lb, dp(AA), dp(BB), dp(c), ip(si), ip(sj), 4, dp(S), nao, ctypes.c_bool(True)) == 0
                ch = np.ascontiguousarray(rng.uniform(-2, 2, (3, 3)))
                V = np.zeros((nao, nao, 3))
                assert lib.nuclear_attraction(la, lb, dp(P), dp(AA), dp(BB), dp(c), ip(si), ip(sj), 4, dp(ch), 3, dp(V),
                                              nao, ctypes.c_bool(True), ctypes.c_double(omega)) == 0
                tag = f"1e_{la}{lb}_w{omega}"
                out[tag + "_S"] = S[:2 * la + 1, 2 * la + 1:]
                out[tag + "_V"] = V[:2 * la + 1, 2 * la + 1:, :]
                out[tag + "_in"] = np.concatenate([P.ravel(), AA.ravel(), BB.ravel(), c, ch.ravel()])
        for lj in range(10):
            for li in range(lj + 1):
                A = np.ascontiguousarray(np.c_[rng.uniform(-1, 1, (2, 3)), rng.uniform(0.3, 2.5, 2)])
                B = np.ascontiguousarray(np.c_[rng.uniform(-1, 1, (2, 3)), rng.uniform(0.3, 2.5, 2)])
                ci, cj = rng.normal(size=2), rng.normal(size=2)
                ni, nj = 2 * li + 1, 2 * lj + 1
                si = np.zeros(2, np.int32); sj = np.full(2, ni, np.int32)
                J = np.zeros((ni + nj, ni + nj))
                rc = lib.two_center(li, lj, dp(A), dp(ci), ip(si), 2, dp(B), dp(cj), ip(sj), 2, dp(J), ni + nj,
                                    ctypes.c_bool(True), ctypes.c_double(omega))
                assert rc == 0
                tag = f"2c_{li}{lj}_w{omega}"
                out[tag + "_J"] = J[:ni, ni:]
                out[tag + "_in"] = np.concatenate([A.ravel(), ci, B.ravel(), cj])
    return out


HOST_BASIS = {"3-21g": ("H", "O", "F"), "def2-tzvp": ("H", "O"), "def2-svp": ("H", "O", "F"),
              "sto-3g": ("H", "O", "F"), "def2-svp-rifit": ("H", "O", "F")}


def main():
    cases, arrays = extract_cases()
    (HERE / "reference_cases.json").write_text(json.dumps(cases, indent=1))
    blocks = {name: element_blocks((BASIS_DIR / (name + ".gaussian.basis")).read_text(), set(els))
              for name, els in HOST_BASIS.items()}
    (HERE / "basis_blocks.json").write_text(json.dumps(blocks, indent=1))
    np.savez_compressed(HERE / "reference_cases.npz", **arrays)
    print(f"{len(cases)} reference cases")
    if "--no-vectors" in sys.argv:
        return
    vec = class_vectors()
    np.savez_compressed(HERE / "ref_class_vectors.npz", **vec)
    print(f"{len(vec)} class vectors")


if __name__ == "__main__":
    main()
