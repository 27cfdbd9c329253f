#!/usr/bin/env python3
"""Extract the generalized Clebsch-Gordan (gtinv) table into a plain data file.

The reference compiles this table into its binary as three C++ initializer
lists in ``src/USER-MLIP/mlip_gtinv_data.cpp`` (``l_array_all`` :37,
``coeffs_all`` :187, ``m_array_all`` :12385).  The coefficients are data, not
code: a potential file's regression coefficients are only meaningful against
exactly these invariants, so the table has to be carried over value for value.
This script text-parses the initializer lists (nothing is compiled or run) and
writes ``gtinv_table.txt`` next to itself, keeping every coefficient as the
exact decimal string of the reference so that ``strtod`` yields the same
doubles the reference's compiler produced.

Output format (whitespace separated, one irrep per header line)::

    gtinv_table <n_irreps>
    irrep <index> <order q> <l_1 ... l_q> <n_terms T>
    <m_1 ... m_q> <coefficient>        (T lines)

Usage: python3 extract_gtinv.py [path/to/mlip_gtinv_data.cpp]
"""
import os
import re
import sys

DEFAULT_SRC = "/root/reference/src/USER-MLIP/mlip_gtinv_data.cpp"


def _section(text, name):
    start = text.index(name + " =")
    body_start = text.index("{", start)
    depth = 0
    for pos in range(body_start, len(text)):
        ch = text[pos]
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
            if depth == 0:
                return text[body_start:pos + 1]
    raise ValueError("unterminated initializer for " + name)


def _nested(body):
    """Parse a brace-nested list of numeric tokens, keeping tokens as strings."""
    tokens = re.findall(r"[{}]|[-+0-9.eE]+", body)
    stack = [[]]
    for tok in tokens:
        if tok == "{":
            stack.append([])
        elif tok == "}":
            done = stack.pop()
            stack[-1].append(done)
        else:
            stack[-1].append(tok)
    assert len(stack) == 1 and len(stack[0]) == 1
    return stack[0][0]


def extract(src_path):
    text = open(src_path).read()
    l_all = _nested(_section(text, "l_array_all"))
    c_all = _nested(_section(text, "coeffs_all"))
    m_all = _nested(_section(text, "m_array_all"))
    if not (len(l_all) == len(c_all) == len(m_all)):
        raise ValueError("table sections disagree in length")
    irreps = []
    for idx, (ls, cs, ms) in enumerate(zip(l_all, c_all, m_all)):
        ls = [int(v) for v in ls]
        if len(cs) != len(ms):
            raise ValueError("irrep %d: %d coefficients but %d m-tuples" % (idx, len(cs), len(ms)))
        terms = []
        for mt, c in zip(ms, cs):
            mt = [int(v) for v in mt]
            if len(mt) != len(ls):
                raise ValueError("irrep %d: m-tuple length mismatch" % idx)
            terms.append((mt, c))
        irreps.append((ls, terms))
    return irreps


def write_table(irreps, out_path):
    with open(out_path, "w") as fh:
        fh.write("gtinv_table %d\n" % len(irreps))
        for idx, (ls, terms) in enumerate(irreps):
            fh.write("irrep %d %d %s %d\n" % (idx, len(ls), " ".join(map(str, ls)), len(terms)))
            for mt, c in terms:
                fh.write("%s %s\n" % (" ".join(map(str, mt)), c))


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gtinv_table.txt")
    irreps = extract(src)
    write_table(irreps, out)
    print("wrote %d irreps, %d terms -> %s" % (len(irreps), sum(len(t) for _, t in irreps), out))


if __name__ == "__main__":
    main()
