import sys, os
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import v3dsfmtoolkit_amd as v
n = 8192
d = v.synth_pair(v.synth_seed(1, 0), n, n, 0.4, 0.05)
m = v.SIFTMatching(0)
m.enableRoundingWarning(False)
m.allocate(n, n)
m.setLeftFeatures(n, d["L"]); m.setRightFeatures(n, d["R"])
m.setKernel(v.KERNEL_I8)
whole = [i for i in range(len(v.scan_variants())) if v.scan_variant_whole(i)][0]
m.setScanVariant(whole)
rule = (float(sys.argv[1]), 0.8)
for _ in range(30): m.findAsync(*rule)
m.synchronize()
m.close()
print("done", rule, flush=True)
