"""The one hardware target: AMD Instinct MI355X (gfx950, CDNA4).

Only the facts the emitter and the thread policy need. The reference keeps a multi-vendor YAML
table (``tensorforge/common/vm/hw_descr_db.yml:48-74``); gfx950 is the only row kept here.
"""


class Gfx950:
  arch = 'gfx950'
  backend = 'hip'
  manufacturer = 'amd'
  vec_unit_length = 64            # wavefront size
  hw_fp_word_size = 4             # bytes per VGPR lane
  mem_access_align_size = 32
  max_local_mem_size_per_block = 160 * 1024
  max_threads_per_block = 1024
  max_reg_per_block = 512 * 1024
  max_threads_per_sm = 2560
  max_block_per_sm = 40
  shmem_banks = 32                # bank count the padding heuristics use (reference table)
  num_cus = 256
  num_xcds = 8
  # published peaks used as roofline denominators
  hbm_bytes_per_s = 8.0e12
  fp64_flops = 78.6e12
  fp32_flops = 157.3e12
  bf16_mfma_flops = 2.5e15


SUPPORTED_ARCHS = {'gfx950': Gfx950}
# 'hip': the gfx950 kernels; 'omptarget': the same chains as host C++/OpenMP launchers (the CPU
# form the reference's omptarget backend targets, used for the CPU baseline and parity)
SUPPORTED_BACKENDS = ('hip', 'omptarget')


def hw_descr_factory(arch, backend):
  """Return the hardware description; anything but gfx950 with the hip or omptarget backend
  raises ``ValueError`` like the reference's factory does for an unknown architecture
  (``hw_descr.py:81``)."""
  if backend not in SUPPORTED_BACKENDS:
    raise ValueError(f'Unknown backend, given: {backend} (gfx950 HIP kernels and their '
                     'omptarget CPU form are emitted)')
  if arch not in SUPPORTED_ARCHS:
    raise ValueError(f'Unknown gpu architecture: {backend} {arch} (only gfx950 is targeted)')
  return SUPPORTED_ARCHS[arch]
