// FP64 instantiation of the step driver.  Built with -ffp-contract=off so the
// kernels perform the reference's double-precision operations one by one.
#include "vgpu_run.h"
template int vgpu::run_steps<double>(vgpu::Ctx *, uint64_t, hipStream_t, bool, const vgpu::VisitRun *);
template int vgpu::prepare_sweeps<double>(vgpu::Ctx *);
template int vgpu::upload_pack<double>(vgpu::Ctx *);
template int vgpu::run_group_steps<double>(vgpu::Ctx *const *, int, uint64_t, hipStream_t);
template int vgpu::check_limits<double>(vgpu::Ctx *, const int (*)[6], vgpu_limits_report *, unsigned char *, void *,
                                         size_t);
