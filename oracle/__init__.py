"""ORACLE — test infrastructure only (never imported by the product).

`oracle.pushpull` wraps the plain-C restatement in `pushpull_oracle.c`;
`oracle.solver` wraps `solver_oracle.c`; `oracle.ref` loads the reference's
own C++ CPU backend when it has been built into `oracle/_ref/`.
Only tests/, the smoke check and the bench's CPU-baseline leg use them.
"""
