"""Generate gl_table_data.inc: the default Gauss-Legendre table compiled into
libmagnus_hip.so.

Like the reference's build step (src/util/gen_nodes.py:11-47, run by
CMakeLists.txt:29-38 with N = 64), nodes and weights come from
numpy.polynomial.legendre.leggauss mapped to [0, 1] by x -> (x + 1) / 2 and
w -> w / 2, ascending, flattened by order (order o at offset o(o-1)/2).
Values are written as C99 hexadecimal literals, so the table is bit-identical
to the one the reference embeds.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

MAX_ORDER = 64


def render(max_order: int = MAX_ORDER) -> str:
    nodes, weights = [], []
    for order in range(1, max_order + 1):
        x, w = np.polynomial.legendre.leggauss(order)
        nodes.extend(0.5 * (x + 1.0))
        weights.extend(0.5 * w)
    total = len(nodes)

    def block(vals):
        lines = []
        for i in range(0, len(vals), 4):
            lines.append("    " + ", ".join(float(v).hex() for v in vals[i:i + 4]) + ",")
        return "\n".join(lines)

    return (
        "// Generated by gen_gl_table.py (numpy leggauss mapped to [0,1]); do not edit.\n"
        f"static constexpr unsigned long long kGLDefaultMaxOrder = {max_order};\n"
        f"static const double kGLDefaultNodes[{total}] = {{\n{block(nodes)}\n}};\n"
        f"static const double kGLDefaultWeights[{total}] = {{\n{block(weights)}\n}};\n"
    )


def main(out_path: str | None = None) -> Path:
    out = Path(out_path) if out_path else Path(__file__).with_name("gl_table_data.inc")
    text = render()
    if not out.exists() or out.read_text() != text:
        out.write_text(text)
    return out


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else None)
