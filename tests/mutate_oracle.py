"""Plants plausible restatement mistakes in a copy of the oracle and checks that
tests/test_oracle_pins.py catches each one (run by hand: ``python tests/mutate_oracle.py``).

The repository itself is never modified: oracle/, tests/ and sparse_wf_amd/ (Python files
only) are copied to a temporary directory, one mistake is written into the copy of
oracle/fire.py, and the pin tests run there.
"""
from __future__ import annotations

import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent

MUTATIONS = {
    "GLU halves swapped": ("return g2[:, :F] * fl.silu(g2[:, F:])", "return g2[:, F:] * fl.silu(g2[:, :F])"),
    "e-e Dense_1 bias dropped": ('e = fl.linear(_log_scaled(feat, d), q["Dense_1"]["kernel"], q["Dense_1"]["bias"])',
                                 'e = fl.linear(_log_scaled(feat, d), q["Dense_1"]["kernel"])'),
    "attention normalised over both spins": ("norm = torch.stack([att[:n_up].sum(0), att[n_up:].sum(0)])",
                                             "norm = torch.stack([att.sum(0), att.sum(0)])"),
    "update MLP without the final silu": ('x = fl.silu(fl.linear(x, up["Linear_1"]["kernel"], up["Linear_1"]["bias"]))\n    h_mlp = x',
                                          'x = fl.linear(x, up["Linear_1"]["kernel"], up["Linear_1"]["bias"])\n    h_mlp = x'),
    "static cutoff in the e-e envelope": ("gamma, e = ee_edge(edge_p, cut, r[ci] - r[ni])\n        msg = gamma",
                                          "gamma, e = ee_edge(edge_p, cfg.cutoff, r[ci] - r[ni])\n        msg = gamma"),
    "LayerNorm eps 1e-5": ("def layer_norm(h, scale, bias, eps=1e-6):", "def layer_norm(h, scale, bias, eps=1e-5):"),
    "features ordered [diff, d]": ("return torch.cat([d.unsqueeze(-1), diff], -1), d",
                                   "return torch.cat([diff, d.unsqueeze(-1)], -1), d"),
    "message split order": ("return [out[:, k * F:(k + 1) * F] for k in range(5)]",
                            "return [out[:, k * F:(k + 1) * F] for k in (0, 1, 3, 2, 4)]"),
    "centre/neighbour messages swapped": ("msg = gamma * fl.silu(e + mct[ci] + mnb[ni])\n        h = h.index_add",
                                          "msg = gamma * fl.silu(e + mct[ni] + mnb[ci])\n        h = h.index_add"),
    "cusp alpha instead of alpha^2": ("term = fl.reciprocal(d + a) * (a * a)", "term = fl.reciprocal(d + a) * a"),
    "Jastrow MLP without the final silu": ("x = fl.silu(x @ W + b)  # SparseMLP", "x = (x @ W + b)  # SparseMLP"),
    "envelope layout (orb, det)": ('out = torch.einsum("nmde,mdeo->ndo", env, pi)', 'out = torch.einsum("nmde,mdeo->nod", env, pi)'),
    "spin swap rotated the wrong way": ("phi = torch.cat([phi[..., n_up:], phi[..., :n_up]], -1)",
                                        "phi = torch.cat([phi[..., -n_up:], phi[..., :-n_up]], -1)"),
    "readout bias on the wrong output": ('zb = torch.cat([torch.zeros(1, dtype=a["bias"].dtype), a["bias"]])',
                                         'zb = torch.cat([a["bias"], torch.zeros(1, dtype=a["bias"].dtype)])'),
}


def main() -> int:
    missed = []
    with tempfile.TemporaryDirectory() as tmp:
        tmp = Path(tmp)
        for d in ("oracle", "tests"):
            shutil.copytree(ROOT / d, tmp / d, ignore=shutil.ignore_patterns("__pycache__", "golden"))
        shutil.copytree(ROOT / "sparse_wf_amd", tmp / "sparse_wf_amd",
                        ignore=shutil.ignore_patterns("__pycache__", "csrc", "*.so"))
        src = (ROOT / "oracle" / "fire.py").read_text()
        for name, (a, b) in MUTATIONS.items():
            assert a in src, f"mutation site not found: {name}"
            (tmp / "oracle" / "fire.py").write_text(src.replace(a, b, 1))
            r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                                "tests/test_oracle_pins.py"], cwd=tmp, capture_output=True, text=True)
            caught = r.returncode != 0
            print(f"{'caught' if caught else 'MISSED'}  {name}", flush=True)
            if not caught:
                missed.append(name)
    return 1 if missed else 0


if __name__ == "__main__":
    sys.exit(main())
