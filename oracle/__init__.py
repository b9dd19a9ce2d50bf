"""FP64 CPU oracle for the DROID-style SLAM inner loop.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``vipe_amd`` product path imports,
calls or links anything in this package.  It is imported only by ``tests/``,
by ``__graft_entry__.smoke()`` (as the checker) and by the ``cpu_baseline``
leg of ``bench.py``.

Every function restates, from scratch, the arithmetic of one reference
function and cites the reference file:line it follows (paths relative to the
nv-tlabs/vipe source tree).  How each function is pinned independently of
its own formula (closed forms, group axioms, finite differences, brute-force
dense solves, ``torch.nn.functional`` equivalents) is listed in the module
docstrings and exercised by ``tests/test_oracle_*.py``.

Modules
-------
lie   : SO3 / RxSO3 / SE3 / Sim3 group maths and the 13 lietorch ops with
        their backward formulas (csrc/lietorch_ext/*.h, lietorch_gpu.cu).
corr  : all-pairs correlation volume, avg-pool pyramid, corr_index
        forward/backward, altcorr forward/backward/indexed
        (vipe/slam/networks/droid_net.py, csrc/droid_net_ext/*.cu).
ba    : slam_ext.ba / ba_extended dense bundle adjustment with the reference
        bookkeeping (csrc/slam_ext/geom_kernels.cu:1626-1847) and a
        brute-force dense Gauss-Newton solve used to pin it.
geom  : slam_ext geometry siblings iproj, projmap, frame_distance,
        depth_filter (csrc/slam_ext/geom_kernels.cu:526-953).
spatial_corr : corr_ext spatial correlation sampler forward/backward
        (csrc/corr_ext/correlation_cuda_kernel.cu).
scatter : scatter_ext sum / mul / mean / min / max with arg_out
        (csrc/scatter_ext/cuda/scatter_cuda.cu, scatter.cpp).
robust : robust kernels of the BA solver (vipe/slam/ba/kernel.py).
graph_step : factor-graph update glue: reprojection, motion features,
        target / weight epilogue (vipe/slam/components/factor_graph.py).
update_op : the DroidNet update operator, ConvGRU + GraphAgg + heads
        (vipe/slam/networks/droid_net.py:393-519), in float64.
"""
