"""Shared helpers: the transcribed reference known answers and builders that
turn their op descriptions into operators of either engine."""
import cmath
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
FRAC_1_SQRT_2 = 0.7071067811865476


def load_known_answers():
    with open(os.path.join(HERE, "golden", "reference_known_answers.json")) as f:
        return json.load(f)


def symbol_table(angle):
    c, s = math.cos(0.5 * angle), math.sin(0.5 * angle)
    r = FRAC_1_SQRT_2
    base = {
        "O": complex(0.0, 0.0), "I": complex(1.0, 0.0), "i": complex(0.0, 1.0),
        "c": complex(c, 0.0), "s": complex(s, 0.0), "is": complex(0.0, s),
        "e": complex(c, s), "e*": complex(c, -s),
        "r": complex(r, 0.0), "ri": complex(0.0, r),
        "q": complex(0.5, 0.5), "q*": complex(0.5, -0.5), "t*": complex(r, -r), "h": complex(0.5, 0.0),
    }
    table = dict(base)
    for k, v in base.items():
        table["-" + k] = -v  # Neg of a num-complex value negates both parts
    return table


def sym(x, table):
    return table[x]


def build_ops(desc, ns, table, angle, gate_cls=None):
    """desc: list of [ctor, *args, {opts}] -> one operator of namespace ns."""
    out = ns.id()
    for item in desc:
        item = list(item)
        opts = item.pop() if item and isinstance(item[-1], dict) else {}
        ctor, args = item[0], item[1:]
        args = [angle if a == "angle" else a for a in args]
        if ctor == "id":
            o = ns.id()
        elif ctor == "matrix1":
            o = ns.matrix1(args[0], [sym(v, table) for v in args[1]])
        elif ctor == "matrix2":
            o = ns.matrix2(args[0], args[1], [sym(v, table) for v in args[2]])
        elif ctor == "raw_h2":
            o = ns.raw_h2(args[0], args[1])
        else:
            o = getattr(ns, ctor)(*args)
        if o is None:
            return None
        if "c" in opts:
            o = o.c(opts["c"])
            if o is None:
                return None
        out = out * o
    return out
