"""FP64 CPU oracle for the two-site DMRG hot path.

TEST INFRASTRUCTURE ONLY.  This package is a numpy restatement of the
reference's two-site DMRG inner loop (algs/DMRG.jl, algs/optimizeMPS.jl,
lib/MPO_MPS/{Lupdate,Rupdate,expect}.jl, lib/Env/makeEnv.jl, lib/models.jl,
lib/MPO/makeMPO.jl, algs/time.jl).  It is imported only by the tests, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``,
always as what GPU results are compared against, never as the thing measured
or shipped (the CPU baseline times it as the CPU reference point).  The product
path (the HIP library behind include/dmrj.h) never imports it.

Pinning (what each part is checked against, see tests/test_oracle_pins.py):
  * models / makeMPO / fullH: exact diagonalisation energy of the open L=10
    Heisenberg chain, -4.258035207282883 (test/MPtask_test.jl:323), and the
    Bethe-free closed forms of tiny chains (L=2: -3/4, L=3: -1).
  * Lupdate / Rupdate / make2site / singlesite_update: brute-force einsum over
    the full index set, the identity <A|H_1site A> = expect(psi, mpo)
    (test/MPmeas_test.jl:158-173), env/makeEnv consistency
    (test/MPmeas_test.jl:178-206) and hermiticity of H_eff.
  * simplelanczos: the closed-form 2x2 eigenvalue.
  * fsvd truncation: reconstruction error^2 == truncerr, orthonormality,
    descending order.  The TensorPACK truncation rule itself is a documented
    choice (TensorPACK 0.14.1 is not vendored): parity unpinned beyond it.
  * dmrg sweeps: converge to the ED ground state energy (L=10, L=20) and the
    Hubbard L=10 value -25.380567747816 +- 1e-4 (test/model_test.jl:362-364);
    dense and QN runs agree (test/dmrg_test.jl:82-85).
  * tdvp_twosite / krylovTimeEvol: the Krylov propagator against scipy's
    expm of the same dense effective Hamiltonian, norm conservation.
"""
