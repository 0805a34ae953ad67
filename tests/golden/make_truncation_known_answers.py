"""Write truncation_known_answers.json: the reference's own known-answer cases.

Inputs and expected outputs are transcribed from the reference test suite
(no reference code is imported or run):
  tests/test_truncate/test_truncation.py:7-147
  tests/test_strategy.py:16-114
Expected errors given there as ``np.sum(s[k:] ** 2)`` are evaluated here with a
sequential left-to-right sum (what np.sum does for these short vectors).
"""

import json
from pathlib import Path

NONE, REL_SV, REL_NORM2, ABS_SV = 0, 1, 2, 3
MAXB = 0x7FFFFFFF


def tail_sq(s, k):
    acc = 0.0
    for x in s[k:]:
        acc += x * x
    return acc


cases = []


def add(src, s, method, tol, max_bond, size, err, exact):
    cases.append(dict(source=src, s=s, method=method, tolerance=tol, max_bond=max_bond,
                      expected_size=size, expected_err=err, exact=exact))


s5 = [1.0, 0.2, 0.01, 0.005, 0.0005]
add("test_truncation.py:7-11", s5, NONE, 2.220446049250313e-16, MAXB, 5, 0.0, True)
for m in (REL_NORM2, REL_SV, ABS_SV):
    for tol in (0.0, 1e-3, 1e-5):
        add("test_truncation.py:13-28", [0.0] * 4, m, tol, MAXB, 1, 0.0, True)
asc = [1e-16, 1e-12, 1e-10, 1e-8, 1e-6, 1e-4, 1e-2, 1.0]
for m in (REL_NORM2, REL_SV, ABS_SV):
    add("test_truncation.py:30-44", asc, m, 0.0, MAXB, 8, 0.0, True)
s = [1.0, 0.1, 0.01, 0.001, 0.0001]
for tol, size in ((0.5, 1), (0.05, 2), (0.005, 3), (0.0005, 4), (0.00005, 5)):
    add("test_truncation.py:46-92", s, REL_SV, tol, MAXB, size, tail_sq(s, size), False)
norm_errors = [1.01010100e-02, 1.01010000e-04, 1.01000000e-06, 9.99999994e-09, 0.0]
for (tol, size), e in zip(((0.1, 1), (1e-3, 2), (1e-5, 3), (1e-7, 4), (1e-9, 5)), norm_errors):
    add("test_truncation.py:94-147", s, REL_NORM2, tol, MAXB, size, e, False)
logv = [1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7]
add("test_strategy.py:17-22", logv, NONE, 2.220446049250313e-16, MAXB, 7, 0.0, True)
add("test_strategy.py:24-31", logv, NONE, 1e-3, MAXB, 7, 0.0, True)
add("test_strategy.py:44-54", logv, ABS_SV, 0.5e-4, MAXB, 4, tail_sq(logv, 4), True)
add("test_strategy.py:56-67", logv, ABS_SV, 0.5e-4, 2, 2, tail_sq(logv, 2), True)
add("test_strategy.py:69-77", logv, ABS_SV, 0.0, MAXB, 7, 0.0, True)
add("test_strategy.py:91-101", logv, REL_SV, 0.5e-4, MAXB, 5, tail_sq(logv, 5), True)
add("test_strategy.py:103-114", logv, REL_SV, 0.5e-4, 2, 2, tail_sq(logv, 2), True)

method_changes = [
    dict(source="test_strategy.py:79-87", method=REL_SV, tolerance=0.0, expected_method=ABS_SV),
    dict(source="test_strategy.py:79-87", method=REL_NORM2, tolerance=0.0, expected_method=ABS_SV),
    dict(source="test_strategy.py:79-87", method=NONE, tolerance=0.0, expected_method=NONE),
]

out = Path(__file__).with_name("truncation_known_answers.json")
out.write_text(json.dumps(dict(cases=cases, method_changes=method_changes), indent=1))
print(out, len(cases))
