"""Autotuning of the tuned emitter's schedule: candidate spaces, parallel compilation, timing
on the GPU, a measurement database and the learned cost model that prunes the search."""
