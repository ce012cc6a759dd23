"""Writes ehss_trig_table.h: the constants of the correctly rounded sin/cos in
ehss_trig.h, computed with mpmath at 200 bits.

    python ehss_amd/csrc/gen_trig_table.py      (rewrites the header in place)

Contents:
  kPio2[4]      pi/2 as four 48-bit pieces (n * piece is exact for n <= 31);
  kTab[j]       sin(j/512) and cos(j/512) for j = 0..kTabN-1, each as a
                double-double (hi = RN(v), lo = RN(v - hi));
  Taylor coefficients 1/6 (double-double), 1/120, 1/5040, 1/362880, 1/24,
  1/720, 1/40320 (RN to double).
"""
from __future__ import annotations

from pathlib import Path

import mpmath as mp

mp.mp.prec = 200
TAB_N = 405  # j = 0..404 covers |r| <= pi/4 + slack at step 1/512


def rn(v) -> float:
    return float(mp.mpf(v))  # mpmath -> float rounds to nearest


def dd(v):
    hi = rn(v)
    lo = rn(v - mp.mpf(hi))
    return hi, lo


def chunks(v, bits=48, count=4):
    out = []
    rest = mp.mpf(v)
    for _ in range(count):
        e = int(mp.floor(mp.log(abs(rest), 2)))
        scale = mp.mpf(2) ** (bits - 1 - e)
        piece = mp.floor(rest * scale) / scale  # truncation keeps every piece's sign
        out.append(float(piece))
        assert mp.mpf(out[-1]) == piece
        rest -= piece
    return out


def h(x: float) -> str:
    return float.hex(x)


def main() -> None:
    pio2 = chunks(mp.pi / 2)
    lines = [
        "// Generated by gen_trig_table.py (mpmath, 200 bits); do not edit.",
        "// Constants of the correctly rounded sin/cos in ehss_trig.h.",
        "#pragma once",
        "",
        "namespace ehss_trig {",
        "",
        f"constexpr int kTabN = {TAB_N};",
        "// pi/2 = kPio2[0] + kPio2[1] + kPio2[2] + kPio2[3] (+ < 2^-190), 48 bits each.",
        "EHSS_TRIG_CONST double kPio2[4] = {" + ", ".join(h(p) for p in pio2) + "};",
        f"constexpr double kTwoOverPi = {h(rn(2 / mp.pi))};",
        "// 1/6 as a double-double, then RN(1/k!).",
        f"constexpr double kInv6Hi = {h(dd(mp.mpf(1) / 6)[0])}, kInv6Lo = {h(dd(mp.mpf(1) / 6)[1])};",
    ]
    for k in (5, 7, 9, 4, 6, 8):
        lines.append(f"constexpr double kInvFact{k} = {h(rn(1 / mp.factorial(k)))};")
    lines.append("// kTab[j] = {sin hi, sin lo, cos hi, cos lo} of j/512.")
    lines.append("EHSS_TRIG_CONST double kTab[kTabN][4] = {")
    for j in range(TAB_N):
        a = mp.mpf(j) / 512
        sh, sl = dd(mp.sin(a))
        ch, cl = dd(mp.cos(a))
        lines.append(f"    {{{h(sh)}, {h(sl)}, {h(ch)}, {h(cl)}}},")
    lines.append("};")
    lines.append("")
    lines.append("}  // namespace ehss_trig")
    out = Path(__file__).with_name("ehss_trig_table.h")
    out.write_text("\n".join(lines) + "\n")
    print(f"wrote {out} ({TAB_N} table rows)")


if __name__ == "__main__":
    main()
