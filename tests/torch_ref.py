"""Float64 torch evaluation of FermiNet / Psiformer log|psi| for the tests.

TEST INFRASTRUCTURE.  An independent restatement built from torch library
operations (layer_norm, softmax, linalg.slogdet) on the parameter tree of
jaqmc_amd.params, for one walker; autograd gives its coordinate and
parameter derivatives.  tests/test_oracle_pins.py checks it against the
numpy oracle (value, gradient, Laplacian) and against finite differences of
the oracle along parameter directions.
"""

import numpy as np
import torch

EPS = 1e-5


def _t(a):
    if isinstance(a, torch.Tensor):
        return a
    return torch.as_tensor(np.asarray(a, np.float64))


def _torch_features(x, atoms, rescale):
    d = x[:, None, :] - atoms[None]  # (N, M, 3)
    r = torch.linalg.norm(d, dim=-1)
    if rescale:
        g = torch.log1p(r)
        ae = torch.cat([g[..., None], d * (g / r)[..., None]], -1)
    else:
        ae = torch.cat([r[..., None], d], -1)
    N = x.shape[0]
    dd = x[:, None, :] - x[None, :, :]
    eye = torch.eye(N, dtype=x.dtype)
    ree = torch.linalg.norm(dd + eye[..., None], dim=-1) * (1 - eye)
    ee = torch.cat([ree[..., None], dd], -1)
    return ae.reshape(N, -1), ee, r, ree


def torch_logpsi(spec, p, x, atoms):
    N = spec.n_elec
    n_up, n_dn = spec.nspins
    K = spec.ndets
    chans = [(0, n_up), (n_up, N)] if (n_up > 0 and n_dn > 0) else [(0, N)]
    if spec.kind == "ferminet":
        h1, h2, r_ae, _ = _torch_features(x, atoms, False)
        L = len(p["single"])
        for l in range(L):
            parts = [h1] + [h1[a:b].mean(0, keepdim=True).expand(N, -1) for a, b in chans] + \
                    [h2[a:b].mean(0) for a, b in chans]
            W, b = p["single"][l]
            new = torch.tanh(torch.cat(parts, -1) @ _t(W) + _t(b))
            h1 = (h1 + new) / np.sqrt(2.0) if new.shape == h1.shape else new
            if l < L - 1:
                W2, b2 = p["double"][l]
                new2 = torch.tanh(h2 @ _t(W2) + _t(b2))
                h2 = (h2 + new2) / np.sqrt(2.0) if new2.shape == h2.shape else new2
        h = h1
    else:
        ae, _, r_ae, _ = _torch_features(x, atoms, spec.rescale)
        spins = torch.cat([torch.ones(n_up), -torch.ones(n_dn)]).to(x.dtype)[:, None]
        h = torch.cat([ae, spins], -1) @ _t(p["input"][0]) + _t(p["input"][1])
        D = h.shape[-1]
        H = spec.num_heads
        ln = lambda y, s, b: torch.nn.functional.layer_norm(y, (D,), _t(s), _t(b), EPS)
        for lp in p["layers"]:
            xin = ln(h, lp["ln1_scale"], lp["ln1_bias"]) if spec.layer_norm_mode == "pre" else h
            q = (xin @ _t(lp["wq"]) + _t(lp["bq"])).reshape(N, H, -1).transpose(0, 1)
            k = (xin @ _t(lp["wk"]) + _t(lp["bk"])).reshape(N, H, -1).transpose(0, 1)
            v = (xin @ _t(lp["wv"]) + _t(lp["bv"])).reshape(N, H, -1).transpose(0, 1)
            a = torch.softmax(q @ k.transpose(1, 2) / np.sqrt(q.shape[-1]), dim=-1)
            o = (a @ v).transpose(0, 1).reshape(N, D)
            h = h + o @ _t(lp["wo"]) + _t(lp["bo"])
            if spec.layer_norm_mode == "post":
                h = ln(h, lp["ln1_scale"], lp["ln1_bias"])
            m = ln(h, lp["ln2_scale"], lp["ln2_bias"]) if spec.layer_norm_mode == "pre" else h
            for W, b in lp["mlp"]:
                m = torch.tanh(m @ _t(W) + _t(b))
            h = h + m
            if spec.layer_norm_mode == "post":
                h = ln(h, lp["ln2_scale"], lp["ln2_bias"])
    # orbitals (K, N, N) and envelope
    if len(p["orbital"]) == 2:
        blocks = [h[a:b] @ _t(p["orbital"][c]) for c, (a, b) in enumerate(chans)]
        if "orbital_bias" in p:
            blocks = [blk + _t(p["orbital_bias"][c]) for c, blk in enumerate(blocks)]
        phi = torch.cat(blocks, 0)
        echans = chans
    else:
        phi = h @ _t(p["orbital"][0])
        if "orbital_bias" in p:
            phi = phi + _t(p["orbital_bias"][0])
        echans = [(0, N)]
    phi = phi.reshape(N, K, N).transpose(0, 1)
    envs = []
    for c, (a, b) in enumerate(echans):
        pi, sg = _t(p["envelope"][c][0]), _t(p["envelope"][c][1])  # (N_orb, M, K)
        r = r_ae[a:b][:, None, :, None]  # (n, 1, M, 1)
        ex = -torch.abs(sg * r) if spec.envelope == "abs_isotropic" else -sg * r
        envs.append(torch.sum(pi * torch.exp(ex), dim=2))  # (n, N_orb, K)
    env = torch.cat(envs, 0).permute(2, 0, 1)
    sgn, ld = torch.linalg.slogdet(phi * env)
    m = ld.max()
    lp = torch.log(torch.abs(torch.sum(sgn * torch.exp(ld - m)))) + m
    if spec.kind == "psiformer" and spec.jastrow:
        _, _, _, ree = _torch_features(x, atoms, False)
        spin = np.array([0] * n_up + [1] * n_dn)
        iu = np.triu(np.ones((N, N)), 1)
        same = _t((spin[:, None] == spin[None, :]) * iu)
        anti = _t((spin[:, None] == 0) & (spin[None, :] == 1))
        jp = _t(p["jastrow"])
        ap, aa = jp[0], jp[1]
        lp = lp + torch.sum(same * (-0.25 * ap ** 2 / (ap + ree))) + torch.sum(anti * (-0.5 * aa ** 2 / (aa + ree)))
    return lp




def param_tree_to_torch(p):
    """The parameter tree with every array a float64 leaf tensor requiring grad."""
    if isinstance(p, dict):
        return {k: param_tree_to_torch(v) for k, v in p.items()}
    if isinstance(p, (list, tuple)):
        return type(p)(param_tree_to_torch(v) for v in p)
    return torch.tensor(np.asarray(p, np.float64), requires_grad=True)


def flat_leaves(spec, p):
    """Leaves of a torch parameter tree in the pack order of jaqmc_amd.params."""
    from jaqmc_amd import params as P

    return [a for _, a in P.param_arrays(spec, p)]
