"""collop -- MI355X (gfx950) Landau collision hot path behind the reference's hooks.

The compute lives in ``libcollop.so`` (hand-written HIP for gfx950, C-ABI in
include/collop.h).  This package is its host-side mirror of the reference
interface: the Workspace tables (MantisWrappers.jl), the GPU hook functions
(collision_gpu.jl, projection_gpu.jl) and the device-resident implicit step
(solver.jl, main.jl).
"""
from ._lib import (CollopError, Context, StepParams, StepStats, load, MODEL_LANDAU, MODEL_LB, PREC_F64,  # noqa: F401
                   PREC_F32C, WARM_EULER, WARM_MLP)
from .hooks import (compute_collision_gpu, compute_collision_gpu32, compute_collision_metric_gpu,  # noqa: F401
                    compute_fs_minus_fp_l2_gpu, compute_G_gpu, compute_lb_rhs_gpu, compute_negative_part_l1_gpu,
                    compute_r_and_entropy_gpu, enable_gpu, eval_loggrad_gpu, gmres_dense_gpu, l2_project_gpu, l2_rhs_gpu,
                    mass_solve_gpu)
from .ic import ensemble_markers, initial_markers, markers  # noqa: F401
from .mesh import Workspace, build_workspace, extraction_p2c1  # noqa: F401
from .params import (ANDERSON_F32C, CHOSEN, CONFIGS, DEFECT, SimParameters, anderson_config,  # noqa: F401
                     anderson_description, chosen_config, chosen_description, cluster_config, config, defect_config,
                     parse_overrides)
from .records import (CONS_COLS, STATS_COLS, RunRecord, checkpoint_path, cons_time_at,  # noqa: F401
                      load_checkpoint, migrate_cons_csv, run_recorded, save_checkpoint, truncate_csv_after)
from .sim import (Ensemble, Simulation, StreamEnsemble, pair_mode, profile, set_device_solve, set_fused_maps,  # noqa: F401
                  set_pair_mode, set_pair_tile, set_weight_path)
from .warmstart import WarmStart, glorot_init  # noqa: F401
