"""Kernel records: the attributes of tectosaur's Kernel type
(tectosaur/kernels/__init__.py:3-18) for the elastic kernels on the hot
path.  The formulas themselves live in the HIP kernels."""
import attr


@attr.s(frozen=True)
class Kernel:
    name = attr.ib()
    spatial_dim = attr.ib()
    tensor_dim = attr.ib()
    needs_obsn = attr.ib()
    needs_srcn = attr.ib()
    scale_type = attr.ib()
    sm_power = attr.ib()
    flip_negate = attr.ib()
    surf_curl_obs = attr.ib(default=False)
    surf_curl_src = attr.ib(default=False)
    kid = attr.ib(default=-1)
    # algorithmic flops per point pair, tectosaur tests/test_farfield.py:82-92
    flops_per_pair = attr.ib(default=0)


elasticU = Kernel('elasticU3', 3, 3, False, False, -3, 1, False, kid=0, flops_per_pair=28)
elasticT = Kernel('elasticT3', 3, 3, False, True, -2, 0, True, kid=1, flops_per_pair=63)
elasticA = Kernel('elasticA3', 3, 3, True, False, -2, 0, True, kid=2, flops_per_pair=63)
elasticH = Kernel('elasticH3', 3, 3, True, True, -1, -1, False, kid=3, flops_per_pair=102)
elasticRT = Kernel('elasticRT3', 3, 3, False, True, -2, 0, True, surf_curl_src=True, kid=4)
elasticRA = Kernel('elasticRA3', 3, 3, True, False, -2, 0, True, surf_curl_obs=True, kid=5)
elasticRH = Kernel('elasticRH3', 3, 3, True, True, -1, 1, False, surf_curl_obs=True,
                   surf_curl_src=True, kid=6)

elastic_kernels = {K.name: K for K in (elasticU, elasticT, elasticA, elasticH)}
regularized_elastic_kernels = {K.name: K for K in (elasticRT, elasticRA, elasticRH)}
kernels = dict(elastic_kernels)
kernels.update(regularized_elastic_kernels)
