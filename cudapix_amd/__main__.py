"""Command line front-end: pi(x) on the GPU, one line of JSON per query.

    python -m cudapix_amd X [Y [C]] [--alpha A] [--terms]
    python -m cudapix_amd --batch FILE      (one x per line; '-' reads stdin)

X, Y and C are integer expressions (decimal, 1e16, 10^16, 2**64 + 1, ...),
as the reference's `pix x [y [c]]` takes them (src/main.cpp:48-72); Y = 0 or
absent picks y from the alpha model (or from --alpha), C defaults to 6.  The
reference evaluates its arguments through the shell (popen of `calc` and
`echo $((...))`, src/main.cpp:93-131); here they are parsed as integer
arithmetic only.  Each output line holds x, y, z, c, alpha, pi(x), the wall
time, and with --terms every term of
    pi(x) = S0 + (S1 + U) + V + S3 + pi(y) - 1 - P2.
A batch file runs as one batched launch sequence.
"""
import argparse
import ast
import json
import operator
import sys

_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.FloorDiv: operator.floordiv, ast.Pow: operator.pow, ast.BitXor: operator.pow,
        ast.LShift: operator.lshift, ast.Mod: operator.mod}


def parse_int(text):
    """An integer from an arithmetic expression: + - * // ** ^ (power, as in
    calc) << %, parentheses, and decimal literals with an optional exponent
    (1e16 = 10**16; the mantissa times the power of ten must be an integer).
    Anything else is a ValueError."""
    s = text.strip().replace("^", "**")

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, int) and not isinstance(node.value, bool):
            return node.value
        if isinstance(node, ast.Constant) and isinstance(node.value, float):
            raise ValueError("non-integer literal")
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.UAdd, ast.USub)):
            v = ev(node.operand)
            return v if isinstance(node.op, ast.UAdd) else -v
        if isinstance(node, ast.BinOp) and type(node.op) in _OPS:
            a, b = ev(node.left), ev(node.right)
            if isinstance(node.op, ast.Pow) and (b < 0 or b > 400):
                raise ValueError("exponent out of range")
            return _OPS[type(node.op)](a, b)
        raise ValueError("unsupported expression")

    # decimal exponents: 1e16, 2.5e15 -> exact integers
    out, i = [], 0
    while i < len(s):
        ch = s[i]
        if ch.isdigit() or (ch == "." and i + 1 < len(s) and s[i + 1].isdigit()):
            j = i
            while j < len(s) and (s[j].isdigit() or s[j] == "."):
                j += 1
            if j < len(s) and s[j] in "eE" and j + 1 < len(s) and (s[j + 1].isdigit() or s[j + 1] in "+-"):
                k = j + 1
                if s[k] in "+-":
                    k += 1
                while k < len(s) and s[k].isdigit():
                    k += 1
                mant, exp = s[i:j], int(s[j + 1:k])
                whole, _, frac = mant.partition(".")
                digits = int((whole or "0") + frac)
                e = exp - len(frac)
                if e < 0:
                    q, r = divmod(digits, 10 ** (-e))
                    if r:
                        raise ValueError("not an integer: " + s[i:k])
                    out.append(str(q))
                else:
                    out.append(str(digits * 10 ** e))
                i = k
                continue
            if "." in s[i:j]:
                raise ValueError("non-integer literal")
            out.append(s[i:j])
            i = j
            continue
        out.append(ch)
        i += 1
    v = ev(ast.parse("".join(out), mode="eval"))
    if not isinstance(v, int):
        raise ValueError("not an integer")
    return v


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m cudapix_amd", description="pi(x) on the GPU (Deleglise-Rivat)")
    ap.add_argument("x", nargs="?", help="x (integer expression)")
    ap.add_argument("y", nargs="?", default="0", help="y (0: from alpha)")
    ap.add_argument("c", nargs="?", default="6", help="wheel primes c, 1..8")
    ap.add_argument("--alpha", type=float, default=0.0, help="y = alpha * cbrt(x) (default: built-in model)")
    ap.add_argument("--terms", action="store_true", help="report every term")
    ap.add_argument("--batch", help="file of x values, one per line ('-': stdin), run as one batch")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    if a.batch is None and a.x is None:
        ap.error("x or --batch is required")
    try:
        y, c = parse_int(a.y), parse_int(a.c)
        if a.batch is not None:
            f = sys.stdin if a.batch == "-" else open(a.batch)
            xs = [parse_int(ln) for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
        else:
            xs = [parse_int(a.x)]
    except (ValueError, SyntaxError) as e:
        ap.error(str(e))
    if any(x < 0 for x in xs) or y < 0 or not 1 <= c <= 8:
        ap.error("x, y must be non-negative and c in 1..8")

    import cudapix_amd as C
    try:
        ctx = C.Context(device=a.device, c=c, alpha=a.alpha)
    except C.CpxError as e:
        print(json.dumps({"error": str(e), "code": e.code}), flush=True)
        return 1
    try:
        if a.batch is not None:
            res = ctx.pi_batch(xs, terms=True)
            for x, t in zip(xs, res):
                print(json.dumps(_record(x, t, a.terms)), flush=True)
        else:
            t = ctx.pi(xs[0], y=y, c=c, terms=True)
            print(json.dumps(_record(xs[0], t, a.terms)), flush=True)
    except C.CpxError as e:
        print(json.dumps({"error": str(e), "code": e.code}), flush=True)
        return 1
    finally:
        ctx.close()
    return 0


def _record(x, t, terms):
    r = {"x": str(x), "y": t["y"], "z": t["z"], "c": t["c"], "alpha": round(t["alpha"], 4), "pi": str(t["pi"]),
         "ms": round(t["ms"], 3)}
    if terms:
        for k in ("ordinary", "trivial", "easy", "hard", "p2", "pi_y"):
            r[k] = str(t[k])
    return r


if __name__ == "__main__":
    sys.exit(main())
