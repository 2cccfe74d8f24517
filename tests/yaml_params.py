"""Translate the reference's YAML parameter files (examples/test/yaml/*.yaml, copied as
fixtures under tests/golden/yaml) into the solver parameter struct, applying the
reference defaults of src/param/xamg_params_details.h:131-236."""
import yaml

ROLE_KEYS = {"solver_params": "solver", "preconditioner_params": "precond",
             "pre_smoother_params": "pre", "post_smoother_params": "post",
             "coarse_grid_solver_params": "coarse"}


def family_defaults(method, list_name):
    """set_family_defaults (xamg_params_details.h:146-236) for the keys the solve uses"""
    d = {"convergence_check": 1 if list_name == "solver" else 0}
    if method in ("Identity", "Direct"):
        return {}
    if method == "Chebyshev":
        return {"convergence_check": 0, "polynomial_order": 2, "spectrum_fraction": 0.3}
    tol = 1e-8 if list_name == "solver" else 0.0
    d.update({"abs_tolerance": tol, "rel_tolerance": tol})
    if method in ("Jacobi", "HSGS"):
        d["max_iters"] = 100 if list_name == "solver" else 2
        d["relax_factor"] = 0.99
    elif method == "MultiGrid":
        d["max_iters"] = 30 if list_name == "solver" else 1
        d.update({"mg_cycle": 0, "mg_reduced_precision": 0, "mg_coarse_matrix_size": 100,
                  "mg_max_levels": 25, "mg_strength_threshold": 0.5, "mg_trunc_factor": 0.25,
                  "mg_Pmax_elements": 4, "mg_max_row_sum": 0.9})
    else:
        d["max_iters"] = 50 if list_name == "solver" else 1
    return d


def load(path):
    with open(path) as f:
        docs = [d for d in yaml.safe_load_all(f) if d]
    cfg = docs[0]
    roles = {}
    for k, name in ROLE_KEYS.items():
        if k in cfg:
            roles[name] = dict(cfg[k])
    if roles.get("precond", {}).get("method") == "MultiGrid":
        for name in ("pre", "post"):
            roles.setdefault(name, {"method": "Jacobi"})
        roles.setdefault("coarse", {"method": "Direct"})
    names = {"solver": "solver", "precond": "preconditioner", "pre": "pre_smoother",
             "post": "post_smoother", "coarse": "coarse_grid_solver"}
    for name, r in roles.items():
        for k, v in family_defaults(r["method"], names[name]).items():
            r.setdefault(k, v)
    return roles


def to_params(X, roles, **extra):
    """pyxamg.Params from a role dictionary"""
    p = X.Params()
    for name in ("solver", "precond", "pre", "post", "coarse"):
        r = roles.get(name)
        if r is None:
            setattr(p, name, X.role(X.NONE))
            continue
        setattr(p, name, X.role(r["method"], max_iters=int(r.get("max_iters", 1)),
                                abs_tol=float(r.get("abs_tolerance", 0.0)),
                                rel_tol=float(r.get("rel_tolerance", 0.0)),
                                conv_check=int(r.get("convergence_check", 0)),
                                relax_factor=float(r.get("relax_factor", 0.99)),
                                poly_order=int(r.get("polynomial_order", 2)),
                                spectrum_fraction=float(r.get("spectrum_fraction", 0.3))))
    pc = roles.get("precond", {})
    p.mg_cycle = int(pc.get("mg_cycle", 0))
    p.mg_prec = 1 if int(pc.get("mg_reduced_precision", 0)) else 0
    p.hsgs_ncores = 1
    p.hsgs_block_rows = 64
    p.coarse_fp64 = 0
    p.index_mode = 0
    p.graph = 1
    for k, v in extra.items():
        setattr(p, k, v)
    return p


def amg_kwargs(roles):
    pc = roles.get("precond", {})
    return {"coarse_size": int(pc.get("mg_coarse_matrix_size", 100)),
            "max_levels": int(pc.get("mg_max_levels", 25)),
            "strength": float(pc.get("mg_strength_threshold", 0.5)),
            "trunc_factor": float(pc.get("mg_trunc_factor", 0.25)),
            "pmax": int(pc.get("mg_Pmax_elements", 4)),
            "max_row_sum": float(pc.get("mg_max_row_sum", 0.9))}
