"""lipika_amd -- MI355X (gfx950) backend for lipika's per-chunk IQ generation.

The chunk body of the reference's ``generate_band`` (src/pipeline.rs:915-983)
runs as hand-written HIP kernels in ``liblpk.so`` behind the C ABI of
``include/lpk.h``; this package is the host-side mirror of the reference's
scene orchestration around it.
"""
from .config import (  # noqa: F401
    AdcConfig, Bpsk, CaCode, CircularKeplerian, Cw, FrontendConfig, GateConfig, NodeClockSpec,
    NonlinearityConfig, Ook, Qam16, ReceiverBandConfig, ReceiverConfig, SceneConfig, SceneMeta,
    SkyConfig, StationaryEcef, StationaryGeodetic, TransmitterConfig,
)
from .pipeline import (  # noqa: F401
    BREAK, Backend, DEFAULT_CHUNK_SAMPLES, ObserverOutput, TimeBase, WindowOptions, WindowSummary,
    prepare, render_window, run_scene,
)
