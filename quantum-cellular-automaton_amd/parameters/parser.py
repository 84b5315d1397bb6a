"""Command-line parameters with the reference's flags (reference parameters/parser.py:14-210).

Every reference flag keeps its name, default and meaning.  Additive options
select the GPU execution: ``--precision`` (fp64 | mixed), ``--device`` and
``--krylov-tol``.  Unlike the reference, parsing happens when ``Parser`` is
constructed (or on the first ``Parser.instance()``), never at import time.
"""
import argparse

from .rules import Rules

INITIAL_STATES = ["blinker", "triple_blinker", "full_blinker", "single", "single_bottom",
                  "all_ket_0", "all_ket_1", "only_outer", "all_ket_1_but_outer",
                  "equal_superposition", "equal_superposition_but_outer", "gradient", "rand"]
FORMATS = ["html", "eps", "jpeg", "jpg", "pdf", "pgf", "png", "ps", "raw", "rgba", "svg", "svgz",
           "tif", "tiff"]


def _positive(value):
    ivalue = int(value)
    if ivalue <= 0:
        raise argparse.ArgumentTypeError("%s is not a positive int value" % value)
    return ivalue


def build_argparser() -> argparse.ArgumentParser:
    """The reference's flags and defaults (reference parameters/parser.py:14-179),
    plus --precision, --device and --krylov-tol."""
    p = argparse.ArgumentParser(description="A classical simulation of the quantum game of life",
                                formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    g = p.add_argument_group("Rules")
    g.add_argument("--num-cells", dest="NUM_CELLS", type=int, default=9,
                   help="The number of cells to use in the simulation.")
    g.add_argument("--distance", dest="DISTANCE", type=int, default=1,
                   help="The distance each cell looks for alive or dead neighbours")
    g.add_argument("--activation-interval", dest="INTERVAL", metavar=("LOWER", "UPPER"), type=int,
                   nargs=2, default=(1, 2),
                   help="Range of alive neighbours required for a flip, upper index is excluded")
    g.add_argument("--num-steps", dest="NUM_STEPS", type=int, default=10000,
                   help="Number of time steps to simulate")
    g = p.add_argument_group("Initial States")
    g.add_argument("--initial-states", dest="INITIAL_STATES", nargs="*", choices=INITIAL_STATES,
                   help="List of initial states")
    g.add_argument("--initial-state-files", dest="INITIAL_STATE_FILES", nargs="*",
                   help="List of .npz files containing mps initial states created by np.savez().")
    g = p.add_argument_group("Algorithm")
    g.add_argument("--algorithm", dest="ALGORITHM", default="exact",
                   choices=["exact", "1tdvp", "2tdvp", "a1tdvp"],
                   help="The algorithm used for the time evolution")
    g.add_argument("--approximative-evolution-method", dest="APPROXIMATIVE_EVOLUTION_METHOD",
                   default="taylor", choices=["taylor", "expm_multiply", "exact_exponential"],
                   help="Evolution method of the a1tdvp bond-dimension estimate")
    g.add_argument("--taylor-steps", dest="TAYLOR_STEPS", type=_positive, default=5,
                   help="Number of terms of the 'taylor' method")
    g.add_argument("--step-size", dest="STEP_SIZE", type=float, default=.005,
                   help="Size of one time step (time step = STEP_SIZE * pi/2)")
    g.add_argument("--max-bond-dim", dest="MAX_BOND_DIM", type=int, default=32,
                   help="Maximum bond dimension of the MPS")
    g.add_argument("--svd-epsilon", dest="SVD_EPSILON", type=float, default=0.00005,
                   help="Truncation threshold on the discarded singular-value norm")
    g.add_argument("--precision", dest="PRECISION", default="fp64", choices=["fp64", "mixed"],
                   help="GPU contraction precision: fp64, or complex64 contractions on FP32 "
                        "matrix cores with FP64 Krylov/SVD (mixed)")
    g.add_argument("--device", dest="DEVICE", type=int, default=0, help="HIP device ordinal")
    g.add_argument("--krylov-tol", dest="KRYLOV_TOL", type=float, default=0.0,
                   help="Krylov exponential tolerance (0 = 1e-14 for fp64, 1e-10 for mixed)")
    g = p.add_argument_group("Plot")
    g.add_argument("--plotting-frequency", dest="PLOTTING_FREQUENCY", type=float, default=1.0,
                   help="Frequency at which time steps are plotted")
    g.add_argument("--plot-sse", dest="PLOT_SSE", action="store_true")
    g.add_argument("--plot-bond-dims", dest="PLOT_BOND_DIMS", action="store_true")
    g.add_argument("--plot-rounded", dest="PLOT_ROUNDED", action="store_true")
    g.add_argument("--plot-classical", dest="PLOT_CLASSICAL", action="store_true")
    g.add_argument("--show", dest="SHOW", action="store_true")
    g.add_argument("--plot-file-path", dest="PLOT_FILE_PATH", default="plots/")
    g.add_argument("--file-formats", dest="FORMATS", nargs="+", default=["pdf"], choices=FORMATS)
    return p


class Parser(object):
    """Parsed parameters; attribute names follow reference parser.py:180-210."""

    def __init__(self, argv=None) -> None:
        args = build_argparser().parse_args(argv)
        self.rules = Rules(ncells=args.NUM_CELLS,
                           activation_interval=range(args.INTERVAL[0], args.INTERVAL[1]),
                           distance=args.DISTANCE, periodic=False)
        self.num_steps = args.NUM_STEPS
        self.step_size = args.STEP_SIZE
        self.algorithm = args.ALGORITHM
        self.approximative_evolution_method = args.APPROXIMATIVE_EVOLUTION_METHOD
        self.taylor_steps = args.TAYLOR_STEPS
        self.plot_sse = args.PLOT_SSE
        self.plot_classical = args.PLOT_CLASSICAL
        self.plot_bond_dims = args.PLOT_BOND_DIMS
        self.plot_rounded = args.PLOT_ROUNDED
        self.show = args.SHOW
        self.plot_frequency = args.PLOTTING_FREQUENCY
        self.plot_file_path = args.PLOT_FILE_PATH
        self.file_formats = args.FORMATS
        self.initial_states = args.INITIAL_STATES or []
        self.initial_state_files = args.INITIAL_STATE_FILES or []
        self.max_bond_dim = args.MAX_BOND_DIM
        self.svd_epsilon = args.SVD_EPSILON
        self.precision = args.PRECISION
        self.device = args.DEVICE
        self.krylov_tol = args.KRYLOV_TOL
        # derived cadence (reference parser.py:207-210)
        self.plot_step_interval = int(1 / (self.plot_frequency * self.step_size))
        self.plot_steps = self.num_steps // self.plot_step_interval
        if self.num_steps % self.plot_step_interval > 0:
            self.plot_steps += 1

    @classmethod
    def instance(cls) -> "Parser":
        """Process-wide parsed arguments (reference parser.py:212-219)."""
        try:
            return cls._instance
        except AttributeError:
            cls._instance = cls()
            return cls._instance
