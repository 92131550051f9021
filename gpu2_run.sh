mkdir -p gpurun_out
timeout -k 10 400 python -u -m pytest tests/test_gpu_parity.py -v --timeout 120 --timeout-method thread > gpurun_out/gpu2_tests.log 2>&1
rc=$?
echo "pytest rc=$rc"
tail -30 gpurun_out/gpu2_tests.log
if [ $rc -ne 0 ] && [ $rc -ne 1 ]; then exit $rc; fi
echo "=== bench"
timeout -k 10 400 python -u bench.py --steps 5 --warmup 2 --plain-steps 2 --cpu-seconds 10 > gpurun_out/gpu2_bench.log 2>&1
rc2=$?
tail -20 gpurun_out/gpu2_bench.log
exit $rc2
