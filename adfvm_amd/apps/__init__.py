"""Case-file entry points: `python -m adfvm_amd.apps.problem <case file>
[orig|perturb|source]` and `python -m adfvm_amd.apps.adjoint <case file>`
(the reference's apps/problem.py and apps/adjoint.py)."""
