"""Write the golden vectors under tests/golden/ from the reference implementation.

Run in the build container only (the reference checkout is not present where
the GPU tests run):

    python tests/golden/make_golden.py /root/reference

It imports the reference package from the given checkout, evaluates it on
small seeded inputs on the CPU, and stores inputs and outputs as ``.npz``
fixtures (plain arrays, no pickles).  The fixtures pin what no closed form
pins: the exact PositionEmbedding initialisation, the reference's fp32 CPU
PCG/CCCP trajectory at the n=2000 configuration, and reference operator
outputs on small random inputs.
"""

import math
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))


def pathloss_problem(n, seed=0, dtype=torch.float32):
    """n locations in [0,100]^2, 3 log-distance sources (exponent 2-4), N(0,1) noise, standardised.

    Draw order on one CPU generator: locations, source positions, exponents,
    noise.  Same as the package's ``laker.data.pathloss_problem``.
    """
    gen = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.rand(n, 2, generator=gen, dtype=dtype) * 100.0
    tx = torch.rand(3, 2, generator=gen, dtype=dtype) * 100.0
    eta = torch.rand(3, generator=gen, dtype=dtype) * 2.0 + 2.0
    y = torch.zeros(n, dtype=dtype)
    for j in range(3):
        d = torch.norm(x - tx[j], dim=1).clamp(min=1.0)
        y = y - 10.0 * eta[j] * torch.log10(d)
    y = y + torch.randn(n, generator=gen, dtype=dtype)
    y = (y - y.mean()) / y.std()
    return x, y


def main(ref_root):
    sys.path.insert(0, ref_root)
    from laker.embeddings import PositionEmbedding
    from laker.kernels import AttentionKernelOperator
    from laker.models import LAKERRegressor
    from laker.preconditioner import CCCPPreconditioner
    from laker.solvers import PreconditionedConjugateGradient

    torch.set_num_threads(8)

    # 1. PositionEmbedding default init (seed 42), d=10 and d=16.
    gen = torch.Generator().manual_seed(123)
    xq = torch.rand(64, 2, generator=gen) * 100.0
    out = {"x": xq.numpy()}
    for d in (10, 16):
        emb = PositionEmbedding(input_dim=2, embedding_dim=d, device=torch.device("cpu"),
                                dtype=torch.float32)
        with torch.no_grad():
            out[f"e{d}"] = emb(xq).numpy()
        out[f"freq{d}"] = emb.freq.numpy()
        out[f"phase{d}"] = emb.phase.numpy()
        out[f"w1_{d}"] = emb.mlp[0].weight.detach().numpy()
        out[f"b1_{d}"] = emb.mlp[0].bias.detach().numpy()
        out[f"w2_{d}"] = emb.mlp[2].weight.detach().numpy()
        out[f"b2_{d}"] = emb.mlp[2].bias.detach().numpy()
    np.savez(os.path.join(HERE, "position_embedding.npz"), **out)

    # 2. Operator outputs on small random inputs (fp32 and fp64, dense and chunked).
    gen = torch.Generator().manual_seed(7)
    rec = {}
    for tag, dt in (("f32", torch.float32), ("f64", torch.float64)):
        e = torch.randn(100, 6, generator=gen, dtype=dt) * 0.6
        x1 = torch.randn(100, generator=gen, dtype=dt)
        x2 = torch.randn(100, 5, generator=gen, dtype=dt)
        q = torch.randn(17, 6, generator=gen, dtype=dt) * 0.6
        op = AttentionKernelOperator(e, lambda_reg=0.05)
        opc = AttentionKernelOperator(e, lambda_reg=0.05, chunk_size=16)
        rec[f"{tag}_e"] = e.numpy()
        rec[f"{tag}_q"] = q.numpy()
        rec[f"{tag}_x1"] = x1.numpy()
        rec[f"{tag}_x2"] = x2.numpy()
        rec[f"{tag}_y1"] = op.matvec(x1).numpy()
        rec[f"{tag}_y2"] = op.matvec(x2).numpy()
        rec[f"{tag}_y1c"] = opc.matvec(x1).numpy()
        rec[f"{tag}_y2c"] = opc.matvec(x2).numpy()
        rec[f"{tag}_diag"] = op.diagonal().numpy()
        rec[f"{tag}_keval"] = op.kernel_eval(q).numpy()
    # clamp regime: large norms so that max ||e||^2 >= 80 (fp32)
    e = torch.randn(64, 4, generator=gen) * 5.0
    x1 = torch.randn(64, generator=gen)
    op = AttentionKernelOperator(e, lambda_reg=0.05)
    rec["clamp_e"] = e.numpy()
    rec["clamp_x1"] = x1.numpy()
    rec["clamp_y1"] = op.matvec(x1).numpy()
    rec["clamp_skip"] = np.array(op.skip_clamp)
    np.savez(os.path.join(HERE, "operator.npz"), **rec)

    # 3. CCCP factors and a PCG solve on a seeded fp64 problem (n=300, N_r=60).
    gen = torch.Generator().manual_seed(11)
    e = torch.randn(300, 6, generator=gen, dtype=torch.float64) * 0.5
    b = torch.randn(300, generator=gen, dtype=torch.float64)
    b2 = torch.randn(300, 4, generator=gen, dtype=torch.float64)
    op = AttentionKernelOperator(e, lambda_reg=1e-2)
    pre = CCCPPreconditioner(num_probes=60, gamma=0.1, max_iter=200, tol=1e-6, verbose=False,
                             device=torch.device("cpu"), dtype=torch.float64)
    pre.build(op.matvec, 300, seed=5)
    gprobe = torch.Generator(device="cpu").manual_seed(5)
    probes = torch.randn(300, 60, generator=gprobe, dtype=torch.float64)
    pcg = PreconditionedConjugateGradient(tol=1e-10, max_iter=300, verbose=False)
    xs = pcg.solve(op.matvec, pre.apply, b)
    pcg2 = PreconditionedConjugateGradient(tol=1e-8, max_iter=300, verbose=False)
    xs2 = pcg2.solve(op.matvec, pre.apply, b2)
    v = torch.randn(300, generator=gen, dtype=torch.float64)
    np.savez(
        os.path.join(HERE, "cccp_pcg_f64.npz"),
        e=e.numpy(), b=b.numpy(), b2=b2.numpy(), probes=probes.numpy(),
        q_basis=pre.q_basis.numpy(), qr_r=pre.qr_r.numpy(),
        isotropic_coef=np.array(pre.isotropic_coef), q_basis_matrix=pre.q_basis_matrix.numpy(),
        q_eigenvalues=pre.q_eigenvalues.numpy(), q_eigenvectors=pre.q_eigenvectors.numpy(),
        v=v.numpy(), pv=pre.apply(v).numpy(),
        x=xs.numpy(), iterations=np.array(pcg.iterations),
        x2=xs2.numpy(), iterations2=np.array(pcg2.iterations),
    )

    # 4. End-to-end fit at the n=2000 configuration (fp32, reference defaults).
    x, y = pathloss_problem(2000, seed=0)
    gen = torch.Generator().manual_seed(1)
    xt = torch.rand(64, 2, generator=gen) * 100.0
    model = LAKERRegressor(embedding_dim=10, lambda_reg=1e-2, gamma=0.1, pcg_tol=1e-6,
                           pcg_max_iter=1000, verbose=False, device="cpu", dtype=torch.float32)
    model.fit(x, y)
    pred = model.predict(xt)
    np.savez(
        os.path.join(HERE, "fit_n2000.npz"),
        x=x.numpy(), y=y.numpy(), xt=xt.numpy(),
        embeddings=model.embeddings.numpy(), alpha=model.alpha.numpy(),
        iterations=np.array(model.pcg_iterations_),
        isotropic_coef=np.array(model.preconditioner.isotropic_coef),
        q_eigenvalues=model.preconditioner.q_eigenvalues.numpy(),
        pred=pred.numpy(), score=np.array(model.score(x, y)),
    )

    # 5. Paper n=3 example (reference tests/test_models.py:37-88) and tiny variance.
    e3 = torch.tensor([[0.241, 0.444], [-0.336, 0.112], [-0.220, 0.353]], dtype=torch.float64)
    y3 = torch.tensor([-66.14, -65.77, -77.30], dtype=torch.float64)
    es = torch.tensor([[0.051, 0.452]], dtype=torch.float64)
    np.savez(os.path.join(HERE, "paper_n3.npz"), e=e3.numpy(), y=y3.numpy(), e_star=es.numpy(),
             lambda_reg=np.array(0.1))
    well_conditioned(ref_root)
    reference_model_file(ref_root)
    print("golden vectors written to", HERE)


def well_conditioned(ref_root):
    """6. The reference's fp32 fit on a well-conditioned instance.

    n=2000 path-loss problem (seed 0) with lambda_reg=100 and pcg_tol=1e-7
    (kappa(lambda I + K) = 24 in FP64): the reference's own fp32 solution is
    within 2e-7 of the FP64 solution on alpha and 4e-6 on predictions, so a
    1e-5 relative agreement with it is a meaningful bound for another fp32
    solve.  kappa is stored with the vectors.
    """
    sys.path.insert(0, ref_root)
    from laker.models import LAKERRegressor

    torch.set_num_threads(8)
    x, y = pathloss_problem(2000, seed=0)
    gen = torch.Generator().manual_seed(1)
    xt = torch.rand(64, 2, generator=gen) * 100.0
    model = LAKERRegressor(embedding_dim=10, lambda_reg=100.0, gamma=0.1, pcg_tol=1e-7,
                           pcg_max_iter=1000, verbose=False, device="cpu", dtype=torch.float32)
    model.fit(x, y)
    pred = model.predict(xt)
    e = model.embeddings.double().numpy()
    a = np.exp(e @ e.T)
    a[np.diag_indices_from(a)] += 100.0
    w = np.linalg.eigvalsh(a)
    np.savez(
        os.path.join(HERE, "fit_n2000_lam100.npz"),
        x=x.numpy(), y=y.numpy(), xt=xt.numpy(), lambda_reg=np.array(100.0), pcg_tol=np.array(1e-7),
        embeddings=model.embeddings.numpy(), alpha=model.alpha.numpy(),
        iterations=np.array(model.pcg_iterations_), pred=pred.numpy(),
        kappa=np.array(w[-1] / w[0]),
    )


def reference_model_file(ref_root):
    """7. A model file written by the reference's ModelPersistence.save.

    n=500 fp32 fit (seed 0, reference defaults) saved with torch.save by the
    reference's own code (laker/persistence.py:74-137), plus its predictions
    at 64 query points, so a loader can be checked against a file it did not
    write.  The file holds tensors, strings, numbers and the embedding
    module's state dict only, so torch.load(weights_only=True) reads it.
    """
    sys.path.insert(0, ref_root)
    from laker.models import LAKERRegressor
    from laker.persistence import ModelPersistence

    torch.set_num_threads(8)
    x, y = pathloss_problem(500, seed=0)
    gen = torch.Generator().manual_seed(2)
    xt = torch.rand(64, 2, generator=gen) * 100.0
    model = LAKERRegressor(embedding_dim=10, lambda_reg=1e-2, gamma=0.1, pcg_tol=1e-6,
                           pcg_max_iter=1000, verbose=False, device="cpu", dtype=torch.float32)
    model.fit(x, y)
    path = os.path.join(HERE, "reference_model_n500.pt")
    ModelPersistence.save(model, path)
    np.savez(os.path.join(HERE, "reference_model_n500_pred.npz"), xt=xt.numpy(),
             pred=model.predict(xt).numpy(), alpha=model.alpha.numpy(), embeddings=model.embeddings.numpy())


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    only = sys.argv[2] if len(sys.argv) > 2 else None
    if only == "well_conditioned":
        well_conditioned(root)
    elif only == "reference_model":
        reference_model_file(root)
    else:
        main(root)
