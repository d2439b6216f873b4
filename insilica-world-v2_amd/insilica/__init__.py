"""insilica (MI355X): the stacked-layer TRiSK dynamical core, FCT transport and the column/radiation pass of InSilica
World V2, with the reference's Python surface (World, Ensemble, Layers, Transport, SWE) over hand-written gfx950 HIP
kernels called through the C ABI of libinsilica_hip.so."""
