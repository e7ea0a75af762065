/* Defaults of the macros this host adds to the reference's config.h set.
   A config.h that does not define them gets the reference behaviour. */
#ifndef PERIODIC
#define PERIODIC 0          /* 1: doubly periodic box without walls      */
#endif
#ifndef PRECISION
#define PRECISION 0         /* 0: FP64, 1: mixed (FP32 pair sums)        */
#endif
#ifndef DIFF_COEF
#define DIFF_COEF 1.0d0     /* factor on the Laplacian, diffusion test   */
#endif
#ifndef DIFF_SOURCE
#define DIFF_SOURCE 1.0d0   /* factor on the manufactured source         */
#endif
#ifndef DT
#define DT 0.0d0            /* > 0: explicit time step                   */
#endif
#ifndef COND_MAX
#define COND_MAX 1.0d4      /* mixed: condition estimate of the FP64 fallback */
#endif
#ifndef JITTER_SEED
#define JITTER_SEED 2026    /* SplitMix64 seed of the POS_PERT jitter    */
#endif
#ifndef HEAT_MODE
#define HEAT_MODE 0         /* 1: diffusion start f = sin 2pi x sin 2pi y */
#endif
#ifndef CELL_SKIN
#define CELL_SKIN 0.08d0    /* cell-list skin / h_eff; 0: a cell build after every pass */
#endif
#ifndef CELL_SPLIT
#define CELL_SPLIT 2        /* cell columns per stencil reach in x (1-4) */
#endif
#ifndef USE_GRAPH
#define USE_GRAPH 1         /* 1: steps replayed from a captured HIP graph */
#endif
