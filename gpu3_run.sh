mkdir -p gpurun_out
echo "=== mfma probe"
timeout -k 10 60 ./_work/probe/mfma_probe
echo "=== rocprof bench"
export TMPDIR=/tmp
timeout -k 10 500 rocprofv3 --kernel-trace --stats -d gpurun_out/prof_bench -o run -- python3 bench.py --steps 3 --warmup 1 --plain-steps 1 --no-cpu > gpurun_out/gpu3_prof.log 2>&1
rc=$?
tail -5 gpurun_out/gpu3_prof.log
find gpurun_out/prof_bench -name "*stats*" | head
exit $rc
