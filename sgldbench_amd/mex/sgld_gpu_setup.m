function sgld_gpu_setup()
%SGLD_GPU_SETUP Builds the device-resident solver context from meshHierarchy_.
%   Call after FEA_SetupVoxelBased (src/FEA_SetupVoxelBased.m:1-28) with the
%   boundary conditions applied (FEA_ApplyBoundaryCondition.m:27-41); sets
%   sgldCtx_, which the GPU branches of the shims in this folder use when
%   GPUsolve_ is true.  The hierarchy is rebuilt on the device from the
%   finest level's voxels, with the level count and first span that
%   Solving_BuildingMeshHierarchy.m:12-23 produced.
	global meshHierarchy_ poissonRatio_ cellSize_ sgldCtx_
	m = meshHierarchy_(1);
	vox = zeros(m.resY, m.resX, m.resZ, 'uint8');
	vox(m.eleMapBack) = 1;
	numLevels = numel(meshHierarchy_);
	nonDyadic = double(numLevels >= 2 && meshHierarchy_(2).spanWidth == 4);
	if ~isempty(sgldCtx_), sgld_ctx_mex('destroy', sgldCtx_); sgldCtx_ = []; end
	% sgld_ctx_create takes numLevels_ / nonDyadic_ as they are before the
	% non-dyadic adjustment (one level fewer, first span 4)
	sgldCtx_ = sgld_ctx_mex('create', vox, numLevels + nonDyadic, nonDyadic, poissonRatio_, cellSize_);
	sgld_ctx_mex('set_fixed', sgldCtx_, logical(m.fixedDOFs));
end
