"""CPU oracle for the optimal-perturb hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package.  The shipped library (twisty_amd) never imports it.
"""
